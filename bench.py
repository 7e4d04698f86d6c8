#!/usr/bin/env python3
"""apply+round benchmark of the matrix-product core on one MI355X.

The metric is apply+round sites/s of the mixed form against the plain FP64
form, with p50/p99 latency, at a stated relative error.  One JSON line:

Headline (`value`, one "step"): the batching workload -- 4096 independent
MPS (N ~ U[8,48], d = 2, chi ~ U{8,16,32,64,128}, Gaussian N(0, 1/chi) cores
drawn on the device) each through one layer of Haar-random real 2-site gates
on the pairs (0,1), (2,3), ... applied as an MPO of bond 4, then TT-rounded
back to its own chi, all chains in lock-step grouped launches, in the mixed
form (contractions on the FP32 matrix cores; Gram matrices, Cholesky, Jacobi
rotations and truncation in FP64), inputs resident in HBM.
value = sites of the batch * steps / elapsed.  Beside it:
  plain_fp64_sites_per_s  the plain FP64 per-chain form (one chain at a time
                          through the same kernels) on a strided sample of the
                          same jobs; speedup_vs_plain_fp64 = value / that;
  batched_fp64_sites_per_s the grouped form in plain FP64;
  error_vs_fp64_oracle    norms and overlaps <Z|X> of checked jobs against
                          the NumPy FP64 oracle (relative to |<Z|X>| and to
                          ||Z|| ||X||), for the mixed and the FP64 batch;
  roofline                the work kernel with the largest share of GPU time
                          (census step with every launch timed by HIP events),
                          its algorithmic rate over the timed steps;
  cpu_baseline            the oracle on a bounded subsample of the jobs.
forms:
  serve_cfg4    a Poisson stream of brickwork-circuit jobs (N ~ U[16,128],
                layers ~ U[4,32], chi_max ~ U{64..512}, cutoff 1e-10, from
                |0...0>) under the continuous-batching server, against the
                per-chain form on the same jobs, p50/p99 latency, oracle
                errors, and a sequential oracle replay on the host;
  chain_cfg1    the chain N=32, d=2, chi=256, D_w=16 (product bond 4096) in
                plain FP64 and mixed, its own roofline, parity and host
                baseline;
  precision_cfg2 the chain N=64, d=4, chi=1024, D_w=8 (product bond 8192),
                cutoff 1e-8: one plain FP64 and one mixed step, the mixed
                error against plain, full-size identities, and the plain form
                against the oracle on a reduced chain;
  api_paths     the drop-in paths beside the headline against the oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"fp64_mfma_tflops": 78.6, "fp32_mfma_tflops": 157.3, "hbm_gbs": 8000.0}
METRIC = "apply+round sites/s (mixed vs plain FP64) + p50/p99 latency at stated rel err"
FORMS = ("batch", "serve", "chain", "precision", "api")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps of the headline batch")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-jobs", type=int, default=4096)
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--plain-seconds", type=float, default=15.0,
                    help="time box of the per-chain plain FP64 sample of the batch")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="time box of each host oracle baseline")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-census", action="store_true",
                    help="time every kernel of the served stream with the library's events (slows it; for the "
                         "served roofline) and report the census in the served form")
    ap.add_argument("--serve-precision", choices=("fp64", "mixed"), default="fp64",
                    help="arithmetic of the served form (the per-chain comparison is always plain FP64); "
                         "FP32 products add ~1e-7 per layer, which long circuits accumulate past 1e-6")
    ap.add_argument("--serve-jacobi-tol", type=float, default=1e-8,
                    help="rotation threshold of the served form's blocked Jacobi (the error budget's, 1e-8 by "
                         "default; 0: the precision's own, 2048 eps in FP64); the per-chain comparison keeps "
                         "the FP64 one")
    ap.add_argument("--serve-plain-seconds", type=float, default=150.0,
                    help="time box of the per-chain comparison of the served form (its complete circuits "
                         "normally end well inside it)")
    ap.add_argument("--chain-steps", type=int, default=5, help="timed steps of the cfg1 chain form")
    ap.add_argument("--only", choices=FORMS, help="run one form only (profiling)")
    ap.add_argument("--graph-steps", type=int, default=3,
                    help="steps of the headline batch replayed from a captured HIP graph (apply + left sweep), "
                         "timed beside the eager steps")
    ap.add_argument("--autotune-load", help="tile table to load instead of tuning (autotune.save format)")
    ap.add_argument("--autotune-save", help="write the tuned tile table here")
    for f in FORMS[1:]:
        ap.add_argument(f"--no-{f}", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision-form", action="store_true", help="same as --only precision")
    ap.add_argument("--cholqr-probe", action="store_true",
                    help="only the copy probe and one CholeskyQR pass on an 8192 x 4096 unfolding (the chain's "
                         "left-sweep shape), for short counter runs")
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
    if args.precision_form:
        args.only = "precision"

    import numpy as np
    import torch

    from syngular_amd import ops

    ops.device()
    if args.copy_probe or args.cholqr_probe:
        copy_probe(np, torch)
    if args.cholqr_probe:
        A = torch.randn(8192, 4096, dtype=torch.float64, device=ops.device()) / 64.0
        ops.cholqr(A)
        torch.cuda.synchronize()
        log("cholqr probe: one CholeskyQR pass on 8192 x 4096")
        return

    def want(f):
        return args.only == f if args.only else not getattr(args, f"no_{f}", False)

    head = batch_headline(np, torch, args) if want("batch") else None
    forms = {}
    torch.cuda.empty_cache()
    # the tile table was tuned on the batching workload's launch classes; the
    # other forms run on the untuned rule, as they do alone
    from syngular_amd import autotune
    autotune.TABLE.clear()
    if want("serve"):
        forms["serve_cfg4"] = serve_form(np, torch, args)
        log(f"serve form: {forms['serve_cfg4']}")
    if want("chain"):
        forms["chain_cfg1"] = chain_form(np, torch, args)
        log(f"chain form: {forms['chain_cfg1']}")
    if want("precision"):
        torch.cuda.empty_cache()      # the 8192-bond sweeps need large contiguous blocks
        forms["precision_cfg2"] = precision_form(np, torch)
        log(f"precision form: {forms['precision_cfg2']}")
    if want("api"):
        forms["api_paths"] = api_form(np, torch)
        log(f"api paths: {forms['api_paths']}")
    if head is None:
        print(json.dumps({"forms": forms}), flush=True)
        return
    later = head.pop("_graph_later", None)
    if later is not None:
        torch.cuda.empty_cache()
        head["graph_form"] = later()
    line = {
        "metric": METRIC,
        "value": head.pop("value"),
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head.pop("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 contractions + f64 Gram/Cholesky/Jacobi/truncation, Jacobi stopped at 1e-8 (mixed)",
        "data": "synthetic (seeded Gaussian cores drawn on the device, Haar-random orthogonal 2-site gates)",
        "config": head.pop("config"),
    }
    line.update(head)
    served = forms.get("serve_cfg4")
    if served:
        line["served_cfg4"] = {k: served[k] for k in ("sites_per_s", "plain_fp64_per_chain_sites_per_s",
                                                     "speedup_vs_plain", "p50_latency_s", "p99_latency_s",
                                                     "precision", "jacobi_tol", "norm2_rel_err_max",
                                                     "overlap_rel_err_max")}
    line["forms"] = forms or None
    print(json.dumps(line), flush=True)


# --------------------------------------------------------------------------
# headline: the batching workload in the mixed form
# --------------------------------------------------------------------------
def batch_headline(np, torch, args):
    from oracle import reference as oref
    from oracle import rounding as ornd
    from syngular_amd import autotune, ops, rounding
    from syngular_amd import batch as B
    from syngular_amd import workloads as W

    t0 = time.perf_counter()
    ns, chis, gates = W.batching_layout(args.batch_jobs, seed=args.batch_seed)
    X = B.ChainBatch.gaussian(ns, chis, 2, seed=args.batch_seed)
    Op = B.gate_layer(ns, np.zeros(len(ns), np.int64), gates)
    torch.cuda.synchronize()
    nsites = int(ns.sum())
    goff = np.concatenate([[0], np.cumsum([W.npairs(int(v), 0) for v in ns])])
    log(f"batch: {len(ns)} chains, {nsites} sites, {X.hbm_bytes() / 2 ** 30:.1f} GiB of cores, "
        f"built in {time.perf_counter() - t0:.1f} s")

    def step(prec="mixed"):
        return B.apply_and_round(X, Op, chis, precision=prec)

    # tile autotuning over the GEMM shape classes this (chi, d, D_w) mix
    # produces: one recording step, then both tiles timed per class
    t1 = time.perf_counter()
    if args.autotune_load:
        autotune.load(args.autotune_load)
        how = f"loaded from {args.autotune_load}"
    else:
        autotune.collect()
        step()
        step("fp64")
        torch.cuda.synchronize()
        autotune.tune()
        how = "tuned"
    if args.autotune_save:
        autotune.save(args.autotune_save)
    tuned = autotune.export()
    log(f"batch autotune: {len(tuned)} shape classes {how} in {time.perf_counter() - t1:.1f} s, "
        f"{sum(e['tile'] == '128x128' for e in tuned)} take 128x128")
    for i in range(args.warmup):
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f"batch warmup {i}: {time.perf_counter() - t1:.3f} s")
    torch.cuda.synchronize()
    ops.profile_begin()
    t1 = time.perf_counter()
    Z = step()
    torch.cuda.synchronize()
    census_ms = 1e3 * (time.perf_counter() - t1)
    census = ops.profile_end()
    dom = dominant_work_kernel(census)
    log(f"batch census step (all kernels timed): {census_ms:.0f} ms; dominant work kernel {dom}")
    del Z
    ops.profile_begin([dom])
    t1 = time.perf_counter()
    for _ in range(args.steps):
        Z = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    timed = ops.profile_end()
    value = nsites * args.steps / elapsed
    ms = 1e3 * elapsed / args.steps
    log(f"batch mixed: {args.steps} steps in {elapsed:.2f} s, {ms:.0f} ms/step, {value:.0f} sites/s")
    roof = kernel_roofline(census, timed, dom, args.steps, BATCH_TRAFFIC)
    # the same step with its apply and left sweep (no host reads in the
    # mixed form) replayed from one captured HIP graph, the right sweep eager;
    # measured after every other form (main calls it last): after a graph
    # has been captured and replayed in this process, later launches of the
    # served stream ran about 20% slower (46.5 -> 57 s), even with the graph
    # reset, so the graph runs once everything else has been timed
    def graph_form():
        if args.graph_steps <= 0:
            return None
        t1 = time.perf_counter()
        cap = B.CapturedApplyLeft(X, Op, mixed=True)
        cap.round(chis)
        torch.cuda.synchronize()
        tc = time.perf_counter() - t1
        t1 = time.perf_counter()
        for _ in range(args.graph_steps):
            cap.round(chis)
        torch.cuda.synchronize()
        gms = 1e3 * (time.perf_counter() - t1) / args.graph_steps
        graph = {"ms_per_step": round(gms, 1), "sites_per_s": round(nsites / (gms * 1e-3), 1),
                 "steps": args.graph_steps, "table_copies": cap.staging.tables, "capture_s": round(tc, 2),
                 "what": "apply + left sweep replayed from one captured HIP graph, right sweep eager"}
        log(f"batch graph form: {graph}")
        cap.close()
        return graph

    kept = [Z.bond[j, 1:int(ns[j])] for j in range(len(ns))]
    textbook = sum(W.apply_round_flops([1] + [int(chis[j])] * (int(ns[j]) - 1) + [1], _gate_bonds(int(ns[j])), 2,
                                       list(kept[j])) for j in range(len(ns)))
    roof["textbook_tflops_per_step"] = round(textbook / (ms * 1e-3) / 1e12, 3)
    roof["textbook_flops_per_step"] = textbook
    log(f"batch dominant kernel {roof['kernel']}: {roof['achieved']} {roof['unit']} ({roof['frac']:.1%}); "
        f"textbook {roof['textbook_tflops_per_step']} TFLOP/s")
    Zm = Z
    # the grouped form in plain FP64
    t1 = time.perf_counter()
    Zp = step("fp64")
    torch.cuda.synchronize()
    fp64_rate = nsites / (time.perf_counter() - t1)
    log(f"batch fp64: {fp64_rate:.0f} sites/s")

    def job_inputs(j):
        xs = X.to_numpy([j])[0]
        return xs, W.gate_mpo_cores(int(ns[j]), 0, gates[goff[j]:goff[j + 1]])

    # per-chain plain FP64 form on a strided sample of the same jobs
    dev = ops.device()
    done, ptime, used = 0, 0.0, []
    for j in range(0, len(ns), max(1, len(ns) // 256)):
        xs, os_ = job_inputs(j)
        Xs = [torch.from_numpy(c).to(dev) for c in xs]
        Os = [torch.from_numpy(c).to(dev) for c in os_]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rounding.apply_and_round(Os, Xs, chi_max=int(chis[j]))
        torch.cuda.synchronize()
        ptime += time.perf_counter() - t1
        done += int(ns[j])
        used.append(j)
        if ptime > args.plain_seconds:
            break
    plain = done / ptime
    log(f"batch: per-chain plain FP64 form {plain:.0f} sites/s on {len(used)} jobs; oracle check next")
    # oracle errors on every 512th job
    check = list(range(0, len(ns), 512))[:8]
    errs = {"mixed": [], "fp64": []}
    outs = {"mixed": Zm.to_numpy(check), "fp64": Zp.to_numpy(check)}
    for q, j in enumerate(check):
        xs, os_ = job_inputs(j)
        ref = ornd.apply_and_round(os_, xs, chi_max=int(chis[j]))[0]
        n2r, ovr, x2 = ornd.norm2(ref), oref.overlap(ref, xs), ornd.norm2(xs)
        for prec in errs:
            c = outs[prec][q]
            ov = oref.overlap(c, xs)
            errs[prec].append((abs(ornd.norm2(c) - n2r) / n2r, abs(oref.overlap(c, ref) - n2r) / n2r,
                               abs(ov - ovr) / abs(ovr), abs(ov - ovr) / np.sqrt(n2r * x2)))
    # overlap_rel_err: <Z|Z_ref> against <Z_ref|Z_ref> (the result's overlap
    # with the oracle's); the overlap with the input <Z|X> is given relative to
    # itself and to ||Z|| ||X|| (a random gate layer leaves |<Z|X>| tiny
    # against the norms, so relative to itself it mostly measures that)
    err = {prec: {"jobs": len(v), "norm2_rel_err_max": max(e[0] for e in v),
                  "overlap_rel_err_max": max(e[1] for e in v),
                  "overlap_with_input_rel_err_max": max(e[2] for e in v),
                  "overlap_with_input_err_over_norms_max": max(e[3] for e in v)} for prec, v in errs.items()}
    log(f"batch oracle errors: {err}")
    cpu = None
    if not args.no_cpu_baseline:
        threads = os.environ.get("OMP_NUM_THREADS") or str(os.cpu_count())
        cdone, t1, cj = 0, time.perf_counter(), []
        for j in range(7, len(ns), 11):
            xs, os_ = job_inputs(j)
            ornd.apply_and_round(os_, xs, chi_max=int(chis[j]))
            cdone += int(ns[j])
            cj.append(j)
            if time.perf_counter() - t1 > args.cpu_seconds:
                break
        dt = time.perf_counter() - t1
        cpu = {"value": round(cdone / dt, 2), "unit": "sites/s", "cores": int(threads), "kind": "port",
               "sample": f"NumPy FP64 oracle, apply + TT-rounding of {len(cj)} jobs of the same batch (every 11th "
                         f"from job 7, {cdone} sites), one after another, {dt:.1f} s"}
        log(f"batch cpu baseline: {cpu}")
    return {"value": round(value, 1), "ms_per_step": round(ms, 1),
            "config": {"workload": f"batching: {len(ns)} independent MPS, N~U[8,48], d=2, chi~U{{8,16,32,64,128}}, "
                                   f"one layer of Haar O(4) 2-site gates (MPO bond 4) on pairs (0,1),(2,3),..., "
                                   f"rounded back to chi; grouped lock-step launches, mixed precision",
                       "jobs": len(ns), "sites": nsites, "phys_dim": 2, "mpo_bond": 4, "form": "batched mixed"},
            "plain_fp64_sites_per_s": round(plain, 1),
            "plain_sample": f"{done} sites of {len(used)} strided jobs, one chain at a time, plain FP64",
            "speedup_vs_plain_fp64": round(value / plain, 2),
            "batched_fp64_sites_per_s": round(fp64_rate, 1),
            "error_vs_fp64_oracle": err,
            "roofline": roof, "cpu_baseline": cpu, "autotune": tuned, "graph_form": None,
            "_graph_later": graph_form,
            "pipeline": {"census_step_ms": round(census_ms, 1), "census_kernels": census_table(census)}}


def _gate_bonds(n):
    """MPO bonds of one gate layer on the pairs (0,1), (2,3), ...: 4 inside a
    pair, 1 across."""
    return [1] + [4 if k % 2 == 1 and k < n - (n % 2) else 1 for k in range(1, n)] + [1]


# --------------------------------------------------------------------------
# cfg1 chain form
# --------------------------------------------------------------------------
def chain_form(np, torch, args):
    from syngular_amd import linalg, ops, rounding
    from syngular_amd import workloads as W

    n, d, chi, dw = 32, 2, 256, 16
    xs, os_ = W.chain_inputs(n, d, chi, dw, seed=2024)
    dev = ops.device()
    X = [torch.from_numpy(x).to(dev) for x in xs]
    O = [torch.from_numpy(o).to(dev) for o in os_]
    log(f"chain: N={n} d={d} chi={chi} D_w={dw} (product bond {chi * dw}), fp64")

    def step(report=None, precision="fp64"):
        return rounding.apply_and_round(O, X, chi_max=chi, cutoff=0.0, report=report, precision=precision)

    step()
    torch.cuda.synchronize()
    ops.profile_begin()
    t0 = time.perf_counter()
    step()
    torch.cuda.synchronize()
    census_ms = 1e3 * (time.perf_counter() - t0)
    census = ops.profile_end()
    dom = dominant_work_kernel(census)
    ops.profile_begin([dom])
    t0 = time.perf_counter()
    for _ in range(args.chain_steps):
        rep = rounding.RoundingReport()
        Z = step(rep)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    timed = ops.profile_end()
    ms = 1e3 * elapsed / args.chain_steps
    rate = n * args.chain_steps / elapsed
    log(f"chain fp64: {ms:.1f} ms/step, {rate:.2f} sites/s")
    roof, pipeline = roofline(census, timed, dom, args.chain_steps, ms, census_ms, W, n, chi, dw, d, rep)
    parity = parity_checks(np, torch, ops, linalg, rounding, W, O, X, Z, rep)
    log(f"chain parity: {parity}")
    del Z, rep
    out = {"workload": f"chain N={n} d={d} chi={chi} D_w={dw} (product bond {chi * dw}), round to {chi}, plain FP64",
           "sites_per_s": round(rate, 3), "ms_per_step": round(ms, 2), "steps": args.chain_steps,
           "roofline": roof, "parity": parity, "pipeline": pipeline}
    out["mixed"] = chain_mixed_form(np, torch, args, O, X, rate)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(np, xs, os_, chi)
    return out


# --------------------------------------------------------------------------
# shared measurement helpers
# --------------------------------------------------------------------------
# Kernels bound by latency (serial rotation rounds in LDS, 64-column
# Cholesky leaves, one-work-group factorisations): no flop or byte count
# describes them, so they get no roofline; bench.py names the largest of
# them beside the roofline when it takes more GPU time than the dominant
# work kernel.
LATENCY_KERNELS = ("chol_diag_kernel", "jacobi_kernel", "qr_kernel", "bjacobi_step_kernel")
# work figures recorded without launches of their own: the bytes the fused
# Jacobi step's update work-groups move (counted on the device in the census
# step only, where every kernel is timed)
WORK_ONLY = {"tjf_step_bytes": "tjf_step_kernel"}


def dominant_work_kernel(prof):
    """The kernel with an algorithmic work count (not in LATENCY_KERNELS)
    that took the most event time."""
    work = {k: v for k, v in prof.items() if k not in LATENCY_KERNELS and k not in WORK_ONLY}
    return max(work.items(), key=lambda kv: kv[1][1])[0]


def _peak(kernel):
    """MI355X dense peak of the unit a work kernel runs on: the FP32 matrix
    cores for the mixed GEMMs, the FP64 matrix cores for the rest."""
    return PEAKS["fp32_mfma_tflops"] if "mixed" in kernel else PEAKS["fp64_mfma_tflops"]


def kernel_roofline(census, timed, dom, steps, traffic_file):
    """Roofline entry of the dominant work kernel: algorithmic work (flops
    for the GEMMs, bytes for the memory-bound kernels) over its event time in
    the timed steps, where only its launches carry events; plus its share of
    the census step's GPU time and the largest latency-bound kernel."""
    total_ms = sum(v[1] for v in census.values())
    dl, dms, dwork = timed[dom]
    for wk, kern in WORK_ONLY.items():
        if kern == dom and wk in census:
            # counted in the census step (one step); the same per step
            dwork = census[wk][2] * steps
    if dom in BYTE_KERNELS:
        achieved = dwork / (dms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAKS["hbm_gbs"], "unit": "GB/s"}
    else:
        achieved = dwork / (dms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": _peak(dom), "unit": "TFLOP/s"}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = measured_traffic(dom, traffic_file)
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
    return roof


def census_table(census):
    def work(k, v):
        for wk, kern in WORK_ONLY.items():
            if kern == k and wk in census:
                return census[wk][2]
        return v[2]
    return {k: {"launches": v[0], "ms": round(v[1], 2),
                "rate": round(work(k, v) / (v[1] * 1e-3) / (1e9 if k in BYTE_KERNELS else 1e12), 3)
                if v[1] else 0.0,
                "rate_unit": "GB/s" if k in BYTE_KERNELS else "TFLOP/s"}
            for k, v in sorted(census.items(), key=lambda kv: -kv[1][1]) if k not in WORK_ONLY}


def roofline(census, timed, dom, steps, ms_per_step, census_ms, W, n, chi, dw, d, rep):
    """The chain's roofline entry (kernel_roofline) and its pipeline: the
    textbook flops of the step and the census step's per-kernel table."""
    roof = kernel_roofline(census, timed, dom, steps, CHAIN_TRAFFIC)
    textbook = W.apply_round_flops(*W.chain_shapes(n, chi, dw), d, list(rep.kept))
    pipeline = {"textbook_flops_per_step": textbook,
                "textbook_tflops": round(textbook / (ms_per_step * 1e-3) / 1e12, 3),
                "census_step_ms": round(census_ms, 1), "census_kernels": census_table(census)}
    return roof, pipeline


# committed rocprofv3 --pmc summaries, per workload (a kernel's bytes per
# launch depend on the problems its launches carry)
BATCH_TRAFFIC = os.path.join(ROOT, "profiles", "batch4096_jacobi_step_traffic.json")
CHAIN_TRAFFIC = os.path.join(ROOT, "profiles", "cholqr_8192x4096_gemm_traffic.json")


def measured_traffic(kernel, path):
    """HBM bytes per launch of `kernel` in this workload from its committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by a known-size
    copy), or None when no counter run covers it.  The counters cannot be
    read from inside this process; the .md beside the summary has the
    commands."""
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    return rec.get("bytes_per_launch", {}).get(kernel)


# kernels whose profiler work figure is bytes moved (the rest: flops)
BYTE_KERNELS = ("apply_kernel", "tjf_step_kernel")


def sval_errors(np, got, sv, kept):
    """Retained singular values per bond against the oracle's: the largest
    per-value error |ds_i| / s_i and the largest |ds_i| / s_max."""
    per, rel_max = 0.0, 0.0
    for k in range(1, len(sv)):
        g = got[k - 1].cpu().numpy()[: kept[k]]
        w = sv[k][: kept[k]]
        per = max(per, float(np.max(np.abs(g - w) / w)))
        rel_max = max(rel_max, float(np.max(np.abs(g - w)) / w[0]))
    return {"svals_per_value_rel_err": per, "svals_err_over_smax": rel_max}


def overlap_errors(np, ov, ovr, n2r, x2):
    """|<Z|X> - <Z|X>_ref| relative to |<Z|X>_ref| and to ||Z|| ||X||."""
    return {"overlap_rel_err": abs(ov - ovr) / abs(ovr), "overlap_err_over_norms": abs(ov - ovr) / np.sqrt(n2r * x2)}


def parity_checks(np, torch, ops, linalg, rounding, W, O, X, Z, rep):
    """(1) Reduced chain (N=12, d=2, chi=64, D_w=8) against the NumPy oracle:
    retained singular values per bond (value by value, and relative to
    s_max), final norm and the overlap <Z|X> (relative to |<Z|X>| and to
    ||Z|| ||X||).  (2) Full size, GPU only: Pythagoras ||OX||^2 - ||Z||^2 =
    sum_k w_k (relative to ||OX||^2)."""
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
    n2 = float(linalg.overlap(cores, cores).item())
    n2r = ornd.norm2(ref)
    ov = float(linalg.overlap(cores, Xs).item())
    ovr = oref.overlap(ref, xs)
    out["reduced_chain"] = "N=12 d=2 chi=64 D_w=8"
    out.update(sval_errors(np, r.svals, sv, kept))
    out["norm2_rel_err"] = abs(n2 - n2r) / n2r
    out.update(overlap_errors(np, ov, ovr, n2r, ornd.norm2(xs)))
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
    rounding.apply_and_round(O, X, chi_max=256, precision="mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.chain_steps):
        rounding.apply_and_round(O, X, chi_max=256, precision="mixed")
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
    value = n * args.chain_steps / dt
    err = {"reduced_chain": "N=12 d=2 chi=64 D_w=8", "norm2_rel_err": abs(n2 - n2r) / n2r,
           "kept_match": r.kept == kept[1:]}
    err.update(overlap_errors(np, ov, ovr, n2r, ornd.norm2(xs)))
    err.update(sval_errors(np, r.svals, sv, kept))
    return {"sites_per_s": round(value, 3), "ms_per_step": round(1e3 * dt / args.chain_steps, 2),
            "speedup_vs_plain": round(value / plain_sites_per_s, 3), "error_vs_fp64_oracle": err}


def precision_form(np, torch):
    """The precision workload: N=64 sites, d=4, chi=1024, MPO bond 8
    (product bond 8192), Gaussian cores as in the headline chain, rounded to
    chi=1024 with relative cutoff 1e-8.  One step in plain FP64 and one in the
    mixed form (products on the FP32 matrix cores; Gram matrices, Cholesky,
    Jacobi and truncation in FP64), inputs resident.  The mixed result is
    compared with the plain FP64 result of the same inputs (norm, overlap,
    retained singular values per bond, value by value).  The plain result is
    checked at full size through identities that need no oracle: Pythagoras
    ||OX||^2 - ||Z||^2 = sum_k w_k (||OX||^2 by transfer matrices over the
    product cores, built one site at a time), right-orthonormal cores 1..N-1
    and the keep rule on every bond's reported spectrum; and against the
    NumPy FP64 oracle on a reduced chain of the same d, MPO bond and cutoff
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
    for prec in ("fp64", "mixed", "bf16"):
        log(f"precision form {prec}: one step of {out['workload']}")
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        rep = rounding.RoundingReport()
        Z = rounding.apply_and_round(O, X, chi_max=chi, cutoff=cut, report=rep, precision=prec)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        reps[prec], res[prec] = rep, Z
        out[prec] = {"seconds_per_step": round(dt, 3), "sites_per_s": round(n / dt, 3)}
        log(f"precision form {prec}: {dt:.1f} s per step")
    zp = res["fp64"]
    n2p = float(linalg.overlap(zp, zp).item())
    for prec in ("mixed", "bf16"):
        zm = res[prec]
        n2m = float(linalg.overlap(zm, zm).item())
        ov = float(linalg.overlap(zm, zp).item())
        per, smax = 0.0, 0.0
        for sp, sm, kp in zip(reps["fp64"].svals, reps[prec].svals, reps["fp64"].kept):
            a, b = sp.cpu().numpy()[:kp], sm.cpu().numpy()[:kp]
            c = min(len(a), len(b))
            per = max(per, float(np.max(np.abs(a[:c] - b[:c]) / a[:c])))
            smax = max(smax, float(np.max(np.abs(a[:c] - b[:c])) / a[0]))
        out[prec]["speedup_vs_fp64"] = round(out["fp64"]["seconds_per_step"] / out[prec]["seconds_per_step"], 3)
        out[prec]["error_vs_fp64_gpu"] = {"norm2_rel_err": abs(n2m - n2p) / n2p,
                                          "overlap_rel_err": abs(ov - n2p) / n2p,
                                          "svals_per_value_rel_err": per, "svals_err_over_smax": smax,
                                          "kept_match": reps["fp64"].kept == reps[prec].kept}
    out["mixed"]["form"] = "all products on the FP32 matrix cores (FP32 accumulation)"
    out["bf16"]["form"] = ("contractions (absorb, C V, U^T C, P M) on the BF16 matrix cores with FP32 "
                           "accumulation, CholeskyQR products on the FP32 ones")
    kept_p = reps["fp64"].kept
    out["kept_bonds_fp64"] = [min(kept_p), max(kept_p)]
    del res, zm
    # full-size identities of the plain result
    log("precision form: full-size identities (Pythagoras, orthonormality, keep rule)")
    e = torch.ones((1, 1), dtype=torch.float64, device=dev)
    for o, x in zip(O, X):
        # E' = sum_s P_s^T E P_s, with P_s = P[:, s, :]: T = E P (l x d r), then
        # E' = (P as (l d) x r)^T (T as (l d) x r)
        P = ops.apply_many([(o, x)])[0]
        l, r = P.shape[0], P.shape[2]
        T = ops.gemm(e, P.reshape(l, -1))
        e = ops.gemm(P.reshape(-1, r).t(), T.reshape(-1, r))
        del P, T
    full2 = float(e.reshape(-1)[0].item())
    wsum = sum(float(t.item()) for t in reps["fp64"].discarded)
    orth = 0.0
    for c in zp[1:]:
        R = c.reshape(c.shape[0], -1)
        G = ops.gemm(R, R.t())
        orth = max(orth, float((G - torch.eye(G.shape[0], dtype=G.dtype, device=G.device)).abs().max().item()))
    keep_ok = all(ornd.keep_count(s.cpu().numpy(), chi, cut) == c for s, c in zip(reps["fp64"].svals, kept_p))
    out["fp64_full_size"] = {"pythagoras_rel_err": abs(full2 - n2p - wsum) / full2,
                             "discarded_weight_rel": wsum / full2, "right_orthonormality_max_err": orth,
                             "keep_rule_holds_every_bond": keep_ok}
    log(f"precision form: {out['fp64_full_size']}")
    del zp, e
    # the plain form against the oracle, reduced chain of the same d and D_w
    xs, os_ = W.chain_inputs(8, d, 64, dw, seed=65)
    Xs = [torch.from_numpy(x).to(dev) for x in xs]
    Os = [torch.from_numpy(o).to(dev) for o in os_]
    r = rounding.RoundingReport()
    cores = rounding.apply_and_round(Os, Xs, chi_max=64, cutoff=cut, report=r)
    ref, sv, kept, _ = ornd.apply_and_round(os_, xs, chi_max=64, cutoff=cut)
    n2, n2r = float(linalg.overlap(cores, cores).item()), ornd.norm2(ref)
    ovg, ovr = float(linalg.overlap(cores, Xs).item()), oref.overlap(ref, xs)
    fo = {"reduced_chain": f"N=8 d={d} chi=64 D_w={dw}", "norm2_rel_err": abs(n2 - n2r) / n2r,
          "kept_match": r.kept == kept[1:]}
    fo.update(overlap_errors(np, ovg, ovr, n2r, ornd.norm2(xs)))
    fo.update(sval_errors(np, r.svals, sv, kept))
    out["fp64_vs_oracle"] = fo
    return out


def serve_form(np, torch, args):
    """A Poisson stream through the continuous-batching server against the
    per-chain form on the same jobs (time-boxed), with latency percentiles
    and the error against the oracle on the cheapest jobs."""
    from syngular_amd import ops, rounding, serving
    from syngular_amd import workloads as W

    jobs = W.serving_jobs(args.serve_jobs, args.serve_rate, seed=4)
    prec = args.serve_precision
    jtol = args.serve_jacobi_tol if args.serve_jacobi_tol > 0 else None
    srv = serving.Server(serving.DeviceExecutor.lanes(precision=prec, jacobi_tol=jtol), hbm_budget=200 * 2 ** 30,
                         keep_states=False, lanes=serving.chi_lanes)
    # warm-up on a tiny stream (kernel loads, allocator)
    serving.Server(serving.DeviceExecutor(), hbm_budget=2 ** 34).run(
        W.serving_jobs(2, 1e6, seed=99, n_range=(16, 20), layer_range=(2, 2), chis=(64,)))
    torch.cuda.synchronize()
    log(f"serve: {len(jobs)} jobs at {args.serve_rate}/s into the server")
    if args.serve_census:
        ops.profile_begin()
    rep = srv.run(jobs)
    torch.cuda.synchronize()
    census = ops.profile_end() if args.serve_census else None
    log(f"serve: {rep.site_layers} site-layers served in {rep.makespan:.1f} s; per-chain sample next")
    lat = rep.latency_percentiles()
    pyth = max(abs(r.norm2 + r.discarded - 1.0) for r in rep.results)
    # per-chain form: complete circuits of every fourth job of the same
    # stream, one chain and one layer at a time, inside a time box (a box that
    # closed early would leave the later, heavier layers out and favour this
    # form; the result says when it did)
    dev = ops.device()
    done, t0, boxed = 0, time.perf_counter(), False
    subset = jobs[::4]
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
    srv2 = serving.Server(serving.DeviceExecutor(precision=prec, jacobi_tol=jtol), hbm_budget=2 ** 36,
                          keep_states=True)
    got = {r.id: r for r in srv2.run([dict(j, arrival=0.0) for j in cheap]).results}
    errs = []
    from oracle import reference as oref
    for j in cheap:
        cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
        for layer in range(j["layers"]):
            mpo = W.gate_mpo_cores(j["n"], layer % 2, W.layer_gates(j["seed"], j["id"], layer, j["n"]))
            cores = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])[0]
        g = got[j["id"]].cores
        n2r = ornd.norm2(cores)
        errs.append((abs(ornd.norm2(g) - n2r) / n2r, abs(oref.overlap(g, cores) - n2r) / n2r))
    cpu = None
    if not args.no_cpu_baseline:
        # the same stream replayed on the host by the oracle, one job and one
        # layer at a time, time-boxed (the early layers of a circuit are the
        # cheap ones, so the box favours the host)
        threads = os.environ.get("OMP_NUM_THREADS") or str(os.cpu_count())
        cdone, t1, cj, stop = 0, time.perf_counter(), [], False
        for j in jobs[::4]:
            cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
            cj.append(j["id"])
            for layer in range(j["layers"]):
                mpo = W.gate_mpo_cores(j["n"], layer % 2, W.layer_gates(j["seed"], j["id"], layer, j["n"]))
                cores = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])[0]
                cdone += j["n"]
                if time.perf_counter() - t1 > args.cpu_seconds:
                    stop = True
                    break
            if stop:
                break
        dt = time.perf_counter() - t1
        cpu = {"value": round(cdone / dt, 2), "unit": "site-layers/s", "cores": int(threads), "kind": "port",
               "sample": f"NumPy FP64 oracle replaying the circuits of jobs {cj} of the same stream (every 4th) "
                         f"one layer at a time, {cdone} site-layers in {dt:.1f} s (stopped at the box: later, "
                         f"larger-bond layers are left out, which favours the host)"}
        log(f"serve cpu baseline: {cpu}")
    return {"workload": f"Poisson stream of {len(jobs)} jobs at {args.serve_rate}/s: N~U[16,128], layers~U[4,32] "
                        f"brickwork Haar O(4) gates (MPO bond 4), chi_max~U{{64,128,256,512}}, cutoff 1e-10, "
                        f"from |0...0>; {prec}" + (f", blocked Jacobi stopped at {jtol:g}" if jtol else "")
                        + " (served, grouped launches)",
            "precision": prec,
            "jacobi_tol": jtol,
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
            "overlap_rel_err_max": max(e[1] for e in errs),
            "roofline": served_roofline(),
            "census": None if census is None else {k: list(v) for k, v in census.items()},
            "cpu_baseline": cpu}


SERVE_PROFILE = os.path.join(ROOT, "profiles", "serve_stream48_kernel_stats.json")


def served_roofline():
    """Roofline of the served form's dominant work kernel, from the committed
    rocprofv3 kernel-trace summary of a served run (the server's lanes run in
    host threads of their own, which the in-process event timer does not
    follow): algorithmic work per launch over the average launch duration."""
    try:
        with open(SERVE_PROFILE) as f:
            return json.load(f)["roofline"]
    except (OSError, ValueError, KeyError):
        return None


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
