"""Tile autotuner of the grouped GEMM (the products of the many-chain sweeps).

A grouped launch runs all its products with one of two tile configurations
of the MFMA GEMM kernel: 64 x 64 (tile 0) or 128 x 128 (tile 1).  The
choice is made per launch (one configuration per launch keeps a launch
whole; splitting a launch by problem shape measured 10% slower on the
served stream, where the launches are many and short).  It is keyed by the
launch's shape class:

    (ceil log2 of the summed output area, ceil log2 of the mean output area
     per problem, ceil log2 of the mean K, compute kind)

The classes a workload produces follow from its (chi, d, D_w) mix: the
absorb products are (r x r)(r x d P), the Gram products (g x g) over l or
d r, the singular-vector products (l x c) over d r, with P = chi D_w.

Untuned, a launch takes the 128 x 128 tile when its output area fills the
chip with such tiles (at least 256 of them) and their padding wastes at most
15% more area than the 64 x 64 tiles'.  ``collect()`` records, per class,
the largest launch the run issues (its problem shapes); ``tune()`` replays
each recorded launch on synthetic operands of the same shapes with both
configurations (HIP events on the current stream) and keeps the faster in
``TABLE``.  A chain's arithmetic can therefore depend on the launch it
shares (its tile), which moves results at the rounding level only.
"""
from __future__ import annotations

import math

import numpy as np

TABLE: dict = {}
_seen: dict = {}
_collecting = [False]
TIMINGS: dict = {}
COUNTS: dict = {}
KINDS = (False, True, "bf16")   # compute kinds: FP64, FP32, BF16 matrix cores
MAX_RECORDED = 16384            # problems kept per recorded launch


def _clog2(x):
    return math.ceil(math.log2(max(float(x), 1.0)))


def launch_class(m, n, k, kind):
    """Shape-class key of one launch (problems m, n, k; kind 0 FP64,
    1 FP32, 2 BF16)."""
    m, n, k = (np.asarray(v, dtype=np.float64) for v in (m, n, k))
    area = m * n
    cnt = max(len(area), 1)
    return ((_clog2(area.sum()) * 64 + _clog2(area.sum() / cnt)) * 64 + _clog2(k.sum() / cnt)) * 4 + int(kind)


def decode(key):
    key = int(key)
    kind = KINDS[key & 3]
    key >>= 2
    bk = key % 64
    key //= 64
    ba = key % 64
    bt = key // 64
    return {"area<=": 1 << bt, "mean_area<=": 1 << ba, "mean_k<=": 1 << bk, "mixed": kind}


def default_tile(m, n):
    """Untuned rule (see the module docstring)."""
    m, n = np.asarray(m, dtype=np.int64), np.asarray(n, dtype=np.int64)
    if int((m * n).sum()) < 256 * 128 * 128:
        return 0
    pad64 = (((m + 63) // 64) * ((n + 63) // 64)).sum() * 64 * 64
    pad128 = (((m + 127) // 128) * ((n + 127) // 128)).sum() * 128 * 128
    return 1 if pad128 <= 1.15 * pad64 else 0


def tile(m, n, k, kind):
    """Tile configuration of a launch (table, else rule)."""
    key = launch_class(m, n, k, kind)
    if _collecting[0]:
        area = int((np.asarray(m, np.int64) * np.asarray(n, np.int64)).sum())
        old = _seen.get(key)
        if old is None or area > old[0]:
            sel = slice(0, MAX_RECORDED)
            _seen[key] = (area, np.array(m, np.int64)[sel], np.array(n, np.int64)[sel], np.array(k, np.int64)[sel])
    t = TABLE.get(key)
    return default_tile(m, n) if t is None else t


def collect(on=True):
    """Record the launches that follow, one per class (for tune())."""
    _collecting[0] = bool(on)
    if on:
        _seen.clear()


def tune(max_bytes=4 << 30, reps=3):
    """Replay every recorded launch on synthetic operands with both tile
    configurations and keep the faster in TABLE (problems beyond max_bytes
    of operands are left out of the replay)."""
    import torch

    from . import _lib, ops
    from ._lib import call

    dev = ops.device()
    for key in sorted(_seen):
        _, m, n, k = _seen[key]
        per = 8 * (m * k + k * n + m * n)
        keep = np.cumsum(per) <= max_bytes
        keep[0] = True
        m, n, k, per = m[keep], n[keep], k[keep], per[keep]
        cnt = len(m)
        COUNTS[int(key)] = cnt
        kind = decode(key)["mixed"]
        oa = np.concatenate([[0], np.cumsum(m * k)])
        ob = np.concatenate([[0], np.cumsum(k * n)])
        oc = np.concatenate([[0], np.cumsum(m * n)])
        A = torch.randn(int(oa[-1]), dtype=torch.float64, device=dev)
        Bt = torch.randn(int(ob[-1]), dtype=torch.float64, device=dev)
        C = torch.empty(int(oc[-1]), dtype=torch.float64, device=dev)
        res = {}
        for t in (0, 1):
            bm = 128 if t else 64
            tn, tm = (n + bm - 1) // bm, (m + bm - 1) // bm
            it = np.zeros(cnt, dtype=_lib.GEMM_ITEM)
            it["A"] = np.uint64(A.data_ptr()) + (8 * oa[:-1]).astype(np.uint64)
            it["B"] = np.uint64(Bt.data_ptr()) + (8 * ob[:-1]).astype(np.uint64)
            it["C"] = np.uint64(C.data_ptr()) + (8 * oc[:-1]).astype(np.uint64)
            it["m"], it["n"], it["k"] = m, n, k
            it["lda"], it["ldb"], it["ldc"] = k, n, n
            it["alpha"], it["beta"] = 1.0, 0.0
            st = np.concatenate([[0], np.cumsum(tm * tn)])
            it["block_start"], it["tiles_n"] = st[:-1], tn
            d_it = ops._upload(it)
            fn = ops.gemm_fn(kind)
            total = int(st[-1])
            call(fn, d_it.data_ptr(), cnt, total, t, ops._stream())
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            best = float("inf")
            for _ in range(reps):
                e0.record()
                call(fn, d_it.data_ptr(), cnt, total, t, ops._stream())
                e1.record()
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            res[t] = best
        TABLE[int(key)] = 0 if res[0] <= res[1] else 1
        TIMINGS[int(key)] = res
        del A, Bt, C
    _collecting[0] = False
    return dict(TABLE)


def export():
    """The tuned table: one entry per launch class with both timings (ms)."""
    out = []
    for key, t in sorted(TABLE.items()):
        r = TIMINGS.get(key, {})
        e = decode(key)
        e.update({"problems": COUNTS.get(key), "tile": "128x128" if t else "64x64",
                  "ms_64": round(r.get(0, float("nan")), 4), "ms_128": round(r.get(1, float("nan")), 4)})
        out.append(e)
    return out


def save(path):
    """Write TABLE (with the timings and counts) as JSON."""
    import json
    with open(path, "w") as f:
        json.dump({"table": {str(k): int(v) for k, v in TABLE.items()},
                   "timings": {str(k): {str(t): ms for t, ms in v.items()} for k, v in TIMINGS.items()},
                   "counts": {str(k): int(v) for k, v in COUNTS.items()}}, f)


def load(path):
    """Replace TABLE by one written by save()."""
    import json
    with open(path) as f:
        rec = json.load(f)
    TABLE.clear()
    TABLE.update({int(k): int(v) for k, v in rec["table"].items()})
    TIMINGS.clear()
    TIMINGS.update({int(k): {int(t): ms for t, ms in v.items()} for k, v in rec.get("timings", {}).items()})
    COUNTS.clear()
    COUNTS.update({int(k): int(v) for k, v in rec.get("counts", {}).items()})
