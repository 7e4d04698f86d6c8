"""Tile autotuner of the grouped GEMM (the products of the many-chain sweeps).

Every product of a grouped launch runs with one of two tile configurations
of the MFMA GEMM kernel: 64 x 64 (tile 0) or 128 x 128 (tile 1).  The choice
is made per PROBLEM from its shape class, never from the launch as a whole,
so a chain's arithmetic (the order of its sums) does not depend on which
other chains share its launches; a grouped launch whose problems fall in
both configurations becomes two launches.

Shape class: (ceil log2 m, ceil log2 n, ceil log2 k, mixed); the number of
problems a launch holds is not part of it (that would make a problem's tile
depend on its company), the tuner times each class at the largest count it
saw in one launch.  The classes a workload
produces follow from its (chi, d, D_w) mix: the absorb products are
(r x r) (r x d P), the Gram products (g x g) over l or d r, the
singular-vector products (l x c) over d r, with P = chi D_w the product bond.

Without tuning a problem takes the 128 x 128 tile when both sides span a
full 128-row tile and the padding of the larger tile wastes at most 15% more
area than the smaller one's.  ``collect()`` records the classes a run
launches (with the largest problem count seen per class); ``tune()`` then
times both configurations on synthetic problems of each recorded class
(the largest shape and problem count recorded for the class, HIP events
on the current stream) and keeps the faster one in ``TABLE``.  ``export()`` gives the table
in a readable form.
"""
from __future__ import annotations

import numpy as np

TABLE: dict = {}
_seen: dict = {}
_collecting = [False]
TIMINGS: dict = {}
COUNTS: dict = {}


def _clog2(x):
    x = np.maximum(np.asarray(x, dtype=np.int64), 1)
    return np.ceil(np.log2(x)).astype(np.int64)


def classes(m, n, k, mixed):
    """Shape-class keys (int64) of the problems (m, n, k)."""
    bm, bn, bk = _clog2(m), _clog2(n), _clog2(k)
    return ((bm * 64 + bn) * 64 + bk) * 2 + int(bool(mixed))


def decode(key):
    key = int(key)
    mixed = key & 1
    key >>= 1
    bk = key % 64
    key //= 64
    bn = key % 64
    bm = key // 64
    return {"m": 1 << bm, "n": 1 << bn, "k": 1 << bk, "mixed": bool(mixed)}


def default_tiles(m, n):
    """Untuned rule: 1 (128 x 128) when both sides reach 128 and its padded
    area is within 15% of the 64 x 64 one's, else 0."""
    m, n = np.asarray(m, dtype=np.int64), np.asarray(n, dtype=np.int64)
    pad64 = ((m + 63) // 64) * ((n + 63) // 64) * 64 * 64
    pad128 = ((m + 127) // 128) * ((n + 127) // 128) * 128 * 128
    return ((m >= 128) & (n >= 128) & (pad128 <= 1.15 * pad64)).astype(np.int64)


def tiles(m, n, k, mixed):
    """Tile configuration of every problem of a launch (table, else rule)."""
    keys = classes(m, n, k, mixed)
    out = default_tiles(m, n)
    if _collecting[0]:
        m, n, k = (np.broadcast_to(np.asarray(v, dtype=np.int64), keys.shape) for v in (m, n, k))
        for key in np.unique(keys):
            sel = keys == key
            old = _seen.get(int(key), (0, 0, 0, 0))
            _seen[int(key)] = (max(old[0], int(sel.sum())), max(old[1], int(m[sel].max())),
                               max(old[2], int(n[sel].max())), max(old[3], int(k[sel].max())))
    if TABLE:
        for key in np.unique(keys):
            t = TABLE.get(int(key))
            if t is not None:
                out[keys == key] = t
    return out


def collect(on=True):
    """Record the shape classes of the following launches (for tune())."""
    _collecting[0] = bool(on)
    if on:
        _seen.clear()


def tune(max_bytes=2 << 30, reps=3):
    """Time both tile configurations on every recorded class and keep the
    faster in TABLE.  Synthetic problems take the largest shape and count
    recorded for the class, the count cut to fit max_bytes."""
    import torch

    from . import _lib, ops
    from ._lib import call

    dev = ops.device()
    for key in sorted(_seen):
        count, m, n, k = _seen[key]
        mixed = decode(key)["mixed"]
        per = 8 * (m * k + k * n + m * n)
        cnt = int(max(1, min(count, max_bytes // per)))
        COUNTS[int(key)] = cnt
        A = torch.randn(cnt * m * k, dtype=torch.float64, device=dev)
        B = torch.randn(cnt * k * n, dtype=torch.float64, device=dev)
        C = torch.empty(cnt * m * n, dtype=torch.float64, device=dev)
        idx = np.arange(cnt, dtype=np.uint64)
        res = {}
        for tile in (0, 1):
            bm = 128 if tile else 64
            tn, tm = (n + bm - 1) // bm, (m + bm - 1) // bm
            it = np.zeros(cnt, dtype=_lib.GEMM_ITEM)
            it["A"] = np.uint64(A.data_ptr()) + idx * np.uint64(8 * m * k)
            it["B"] = np.uint64(B.data_ptr()) + idx * np.uint64(8 * k * n)
            it["C"] = np.uint64(C.data_ptr()) + idx * np.uint64(8 * m * n)
            it["m"], it["n"], it["k"] = m, n, k
            it["lda"], it["ldb"], it["ldc"] = k, n, n
            it["alpha"], it["beta"] = 1.0, 0.0
            it["block_start"] = np.arange(cnt, dtype=np.int64) * tm * tn
            it["tiles_n"] = tn
            d_it = ops._upload(it)
            fn = "syn_gemm_mixed_f64" if mixed else "syn_gemm_f64"
            call(fn, d_it.data_ptr(), cnt, cnt * tm * tn, tile, ops._stream())
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            best = float("inf")
            for _ in range(reps):
                e0.record()
                call(fn, d_it.data_ptr(), cnt, cnt * tm * tn, tile, ops._stream())
                e1.record()
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            res[tile] = best
        TABLE[int(key)] = 0 if res[0] <= res[1] else 1
        TIMINGS[int(key)] = res
        del A, B, C
    _collecting[0] = False
    return dict(TABLE)


def export():
    """The tuned table: one entry per class with both timings (ms)."""
    out = []
    for key, t in sorted(TABLE.items()):
        c = decode(key)
        r = TIMINGS.get(key, {})
        out.append({"m<=": c["m"], "n<=": c["n"], "k<=": c["k"], "count": COUNTS.get(key), "mixed": c["mixed"],
                    "tile": "128x128" if t else "64x64",
                    "ms_64": round(r.get(0, float("nan")), 4), "ms_128": round(r.get(1, float("nan")), 4)})
    return out
