"""Continuous-batching server for apply-and-round jobs.

A job is a chain of n sites that starts from |0...0> and goes through
`layers` brickwork layers of two-site gates (pairs start at site layer % 2),
each applied as an MPO and followed by SVD rounding to chi_max with a
relative cutoff.  Jobs arrive over time.  The loop works at layer
granularity:

1. admit every job that has arrived, while the HBM budget and the slot count
   allow it (a job's charge is its peak working set during one layer, see
   ``job_bytes``);
2. run ONE layer for every resident job together: one grouped apply and one
   lock-step rounding over all of them, whatever layer each is at;
3. retire the jobs whose last layer is done (their latency is completion
   minus arrival) and free their charge.

With lanes (``chi_lanes``: by chi_max class) each lane runs this loop for its
own jobs in its own host thread and HIP stream, under the shared budget: a
layer step lasts as long as its heaviest job, so light jobs do not wait for
heavy ones, and their launches fill the compute units that a heavy job's
latency-bound steps leave idle.

So requests that arrive while others are on the device join at the next layer
boundary, their work shares every launch with the jobs already running, and a
job's result does not depend on which others share its launches (each item of
a grouped launch is computed on its own).

The executor is pluggable: ``DeviceExecutor`` runs the gfx950 kernels; the
tests drive the scheduler on the CPU with a stand-in that records calls.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np

# A layer holds the product state (internal bond 4 chi in the pairs: 4x the
# state), the left sweep's factors and absorbed cores, and the right sweep's
# Gram matrices, factors and new cores; 24x the state's bytes bounds it.
WORKING_SET_FACTOR = 24


def job_bytes(n, chi_max, d=2):
    """HBM charge of a job: WORKING_SET_FACTOR * 8 B * n * d * chi_max^2."""
    return WORKING_SET_FACTOR * 8 * n * d * chi_max * chi_max


@dataclass
class JobResult:
    id: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    layers: int = 0
    sites: int = 0
    norm2: float = float("nan")
    discarded: float = 0.0
    cores: list = None

    @property
    def latency(self):
        return self.finished - self.arrival


@dataclass
class ServeReport:
    results: list
    makespan: float
    steps: int
    peak_resident: int
    peak_bytes: int
    site_layers: int
    admitted_during_run: int = 0
    trace: list = field(default_factory=list)

    @property
    def sites_per_s(self):
        return self.site_layers / self.makespan if self.makespan > 0 else float("nan")

    def latency_percentiles(self, qs=(50, 99)):
        lat = np.array([r.latency for r in self.results])
        return {f"p{q}": float(np.percentile(lat, q)) for q in qs}


class _Shared:
    """Admission state shared by the lanes: HBM charge and slots in use."""

    def __init__(self, budget, max_slots):
        self.budget, self.max_slots = budget, max_slots
        self.used = self.resident = self.peak_bytes = self.peak_resident = 0
        self.lock = threading.Lock()

    def try_admit(self, nbytes):
        with self.lock:
            if self.used + nbytes > self.budget or self.resident + 1 > self.max_slots:
                return False
            self.used += nbytes
            self.resident += 1
            self.peak_bytes = max(self.peak_bytes, self.used)
            self.peak_resident = max(self.peak_resident, self.resident)
            return True

    def release(self, nbytes):
        with self.lock:
            self.used -= nbytes
            self.resident -= 1


class Server:
    """Continuous batching over one or more lanes.

    executor: an executor (one lane), or with ``lanes`` a callable
    lane key -> executor (one executor, and for the device one HIP stream,
    per lane).  lanes: None, or a callable job -> lane key; jobs of a lane
    share its layer steps, lanes run concurrently in their own host threads
    and share the HBM budget and the slot count."""

    def __init__(self, executor, hbm_budget, max_slots=4096, clock=time.perf_counter, sleep=time.sleep,
                 keep_states=False, lanes=None, poll=1e-3):
        self.ex = executor
        self.budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self.clock = clock
        self.sleep = sleep
        self.keep_states = keep_states
        self.lanes = lanes
        self.poll = poll

    def run(self, jobs):
        """Serve `jobs` (dicts with id, arrival, n, layers, chi_max, cutoff,
        seed); arrival times are seconds after the call starts."""
        for j in jobs:
            if job_bytes(j["n"], j["chi_max"]) > self.budget:
                raise ValueError(f"job {j['id']} needs {job_bytes(j['n'], j['chi_max'])} B, more than the "
                                 f"whole HBM budget {self.budget} B")
        shared = _Shared(self.budget, self.max_slots)
        results = {}
        t0 = self.clock()
        if self.lanes is None:
            stats = [self._lane(self.ex, list(jobs), shared, results, t0)]
        else:
            groups = {}
            for j in jobs:
                groups.setdefault(self.lanes(j), []).append(j)
            stats, errors = [], []

            def work(key, group):
                try:
                    stats.append(self._lane(self.ex(key), group, shared, results, t0))
                except BaseException as e:     # re-raised in the calling thread
                    errors.append(e)
            threads = [threading.Thread(target=work, args=(k, g), daemon=True) for k, g in sorted(groups.items())]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
            if errors:
                raise errors[0]
        makespan = self.clock() - t0
        trace = sorted(x for st in stats for x in st["trace"])
        return ServeReport([results[j["id"]] for j in jobs], makespan, sum(st["steps"] for st in stats),
                           shared.peak_resident, shared.peak_bytes, sum(st["site_layers"] for st in stats),
                           sum(st["late"] for st in stats), trace)

    def _lane(self, ex, jobs, shared, results, t0):
        pending = sorted(jobs, key=lambda j: (j["arrival"], j["id"]))
        resident = []          # job dicts, same order as the executor state
        done_layers = []       # layers done per resident job
        state = ex.empty()
        steps = site_layers = late = 0
        trace = []
        pi = 0
        while pi < len(pending) or resident:
            now = self.clock() - t0
            new = []
            while (pi < len(pending) and pending[pi]["arrival"] <= now
                   and shared.try_admit(job_bytes(pending[pi]["n"], pending[pi]["chi_max"]))):
                j = pending[pi]
                results[j["id"]] = JobResult(j["id"], j["arrival"], admitted=now)
                new.append(j)
                pi += 1
            if new:
                if resident:
                    late += len(new)
                state = ex.merge(state, ex.init(new))
                resident += new
                done_layers += [0] * len(new)
            if not resident:
                # idle until the next arrival, or until another lane frees room
                wait = pending[pi]["arrival"] - (self.clock() - t0)
                self.sleep(wait if wait > 0 else self.poll)
                continue
            state, stats = ex.layer(state, resident, done_layers)
            steps += 1
            for i, j in enumerate(resident):
                r = results[j["id"]]
                r.discarded += stats["discarded"][i]
                done_layers[i] += 1
                site_layers += j["n"]
            fin = [i for i, j in enumerate(resident) if done_layers[i] >= j["layers"]]
            trace.append((self.clock() - t0, len(resident), len(new), len(fin)))
            if fin:
                out = ex.harvest(state, fin, keep_states=self.keep_states)
                tnow = self.clock() - t0
                for q, i in enumerate(fin):
                    j = resident[i]
                    r = results[j["id"]]
                    r.finished = tnow
                    r.layers = done_layers[i]
                    r.sites = j["n"]
                    r.norm2 = out["norm2"][q]
                    if self.keep_states:
                        r.cores = out["cores"][q]
                    shared.release(job_bytes(j["n"], j["chi_max"]))
                fs = set(fin)
                keepi = [i for i in range(len(resident)) if i not in fs]
                state = ex.select(state, keepi)
                resident = [resident[i] for i in keepi]
                done_layers = [done_layers[i] for i in keepi]
        return {"steps": steps, "site_layers": site_layers, "late": late, "trace": trace}


def chi_lanes(job):
    """Lane key by chi_max class: jobs of similar bond share layer steps."""
    c = job["chi_max"]
    return 0 if c <= 128 else (1 if c <= 256 else 2)


class DeviceExecutor:
    """Runs the layers with the grouped gfx950 kernels (syngular_amd.batch).
    With a stream, every launch of this executor goes to that HIP stream (one
    per lane, so lanes overlap on the device).  jacobi_tol: the blocked
    Jacobi's rotation threshold (None: the precision's own, see
    batch.ChainBatch.tt_round)."""

    def __init__(self, d=2, stream=None, precision="fp64", jacobi_tol=None):
        from . import batch, ops
        self.B = batch
        self.ops = ops
        self.d = d
        self.stream = stream
        self.precision = precision
        self.jacobi_tol = jacobi_tol

    def _ctx(self):
        import contextlib
        import torch
        return torch.cuda.stream(self.stream) if self.stream is not None else contextlib.nullcontext()

    @classmethod
    def lanes(cls, d=2, precision="fp64", jacobi_tol=None):
        """Factory for Server(lanes=...): one executor and one stream per lane."""
        import torch
        return lambda key: cls(d, torch.cuda.Stream(), precision, jacobi_tol)

    def empty(self):
        return self.B.concat([])

    def init(self, jobs):
        with self._ctx():
            return self.B.product_states([j["n"] for j in jobs], self.d)

    def merge(self, a, b):
        return self.B.concat([a, b])

    def select(self, state, idx):
        return state.subset(idx)

    def layer(self, state, jobs, done_layers):
        with self._ctx():
            return self._layer(state, jobs, done_layers)

    def _layer(self, state, jobs, done_layers):
        from .workloads import layer_gates
        offs = np.array([l % 2 for l in done_layers], dtype=np.int64)
        gates = [layer_gates(j["seed"], j["id"], l, j["n"]) for j, l in zip(jobs, done_layers)]
        G = np.concatenate(gates) if gates else np.zeros((0, 4, 4))
        mpo = self.B.gate_layer(state.n, offs, G, self.d)
        recs = []
        chi = np.array([j["chi_max"] for j in jobs], dtype=np.int64)
        cut = np.array([j["cutoff"] for j in jobs], dtype=np.float64)
        new = self.B.apply_and_round(state, mpo, chi, cut, records=recs, precision=self.precision,
                                     jacobi_tol=self.jacobi_tol)
        disc = np.zeros(len(jobs))
        if recs:
            import torch
            allv = torch.cat([r.disc for r in recs]).cpu().numpy()
            o = 0
            for r in recs:
                np.add.at(disc, r.jobs, allv[o:o + len(r.jobs)])
                o += len(r.jobs)
        return new, {"discarded": disc}

    def harvest(self, state, idx, keep_states=False):
        with self._ctx():
            return self._harvest(state, idx, keep_states)

    def _harvest(self, state, idx, keep_states=False):
        sub = state.subset(idx)
        n2 = self.B.overlap(sub, sub).cpu().numpy()
        out = {"norm2": [float(v) for v in n2]}
        if keep_states:
            out["cores"] = sub.to_numpy()
        return out
