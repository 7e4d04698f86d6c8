"""NumPy stand-in for the C ABI, for host-side tests of the grouped path.

Test infrastructure only.  It replaces the library calls of
``syngular_amd.batch`` with NumPy emulations that read and write CPU tensors
at the addresses the item tables carry, and it checks every access of every
item against the allocations the module made (``batch._empty`` /
``batch._h2d`` / item-table uploads are routed to registered CPU tensors).
So a host test can run the whole grouped sweep on the CPU and catch a wrong
leading dimension, a short buffer or a bad block count before any GPU run.

The emulated kernels follow the contracts in include/syngular_amd.h:
GEMM (op(A) op(B) with the transA/transB and leading-dimension rules),
MPO x MPS apply, Householder QR (numpy's QR has the same sign convention),
one-sided Jacobi in Gram mode (x_in = rt^T x_out, rows sorted by decreasing
norm, s = sqrt(row norm)), truncation (the oracle's keep rule), row scaling,
identity and copy.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import torch

from oracle.rounding import keep_count


class BoundsError(AssertionError):
    pass


class Emulator:
    def __init__(self):
        self.regions = []          # (start, end, tensor)
        self.calls = []
        self.checked = True        # bounds-check every access against `regions`

    # ---- memory ------------------------------------------------------------
    def register(self, t):
        start = t.data_ptr()
        self.regions.append((start, start + t.numel() * t.element_size(), t))
        return t

    def _check(self, addr, nbytes, what):
        addr = int(addr)
        if nbytes <= 0 or not self.checked:
            return
        for s, e, _ in self.regions:
            if s <= addr and addr + nbytes <= e:
                return
        raise BoundsError(f"{what}: access [{addr:#x}, +{nbytes}) outside every allocation")

    def buf(self, addr, count, dtype=np.float64, what="buffer"):
        dt = np.dtype(dtype)
        self._check(addr, count * dt.itemsize, what)
        if count == 0:
            return np.zeros(0, dt)
        raw = (ctypes.c_char * (count * dt.itemsize)).from_address(int(addr))
        return np.frombuffer(raw, dtype=dt)

    def mat(self, addr, rows, cols, ld, trans, what):
        """View of op(X) (rows x cols) with X row-major, leading dim ld."""
        if rows == 0 or cols == 0:
            return np.zeros((rows, cols))
        if trans:   # op(X)[i, j] = X[j*ld + i]
            ext = (cols - 1) * ld + rows
            b = self.buf(addr, ext, what=what)
            return np.lib.stride_tricks.as_strided(b, (rows, cols), (8, 8 * ld))
        ext = (rows - 1) * ld + cols
        b = self.buf(addr, ext, what=what)
        return np.lib.stride_tricks.as_strided(b, (rows, cols), (8 * ld, 8))

    def table(self, addr, n, dtype):
        return np.frombuffer(self.buf(addr, n * dtype.itemsize, np.uint8, "item table").tobytes(), dtype=dtype)

    # ---- dispatch ----------------------------------------------------------
    def call(self, name, *args):
        self.calls.append(name)
        fn = getattr(self, name, None)
        if fn is None:
            raise NotImplementedError(f"emulator has no {name}")
        return fn(*args)

    def syn_prof_note_flops(self, fl):
        return 0

    def _blocks_ok(self, it, nel_or_tiles, total, name, per=None):
        nb = nel_or_tiles if per is None else (np.asarray(nel_or_tiles) + per - 1) // per
        st = np.concatenate([[0], np.cumsum(nb)[:-1]]) if len(nb) else np.zeros(0)
        if not np.array_equal(it["block_start"], st) or int(np.sum(nb)) != int(total):
            raise BoundsError(f"{name}: block_start/total_blocks do not match the items")

    def syn_gemm_f64(self, items, n, total, tile, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.GEMM_ITEM)
        bm = 128 if tile else 64
        tiles = ((it["m"] + bm - 1) // bm) * ((it["n"] + bm - 1) // bm)
        self._blocks_ok(it, tiles, total, "gemm")
        if not np.array_equal(it["tiles_n"], (it["n"] + bm - 1) // bm):
            raise BoundsError("gemm: tiles_n")
        for r in it:
            m, nn, k = int(r["m"]), int(r["n"]), int(r["k"])
            A = self.mat(r["A"], m, k, int(r["lda"]), int(r["transA"]), "gemm A")
            Bm = self.mat(r["B"], k, nn, int(r["ldb"]), int(r["transB"]), "gemm B")
            C = self.mat(r["C"], m, nn, int(r["ldc"]), 0, "gemm C")
            res = float(r["alpha"]) * (A @ Bm)
            if float(r["beta"]) != 0.0:
                res = res + float(r["beta"]) * C
            C[...] = res
        return 0

    def syn_gemm_mixed_f64(self, items, n, total, tile, stream):
        """Products of FP32-rounded operands accumulated in FP32."""
        from syngular_amd import _lib
        it = self.table(items, n, _lib.GEMM_ITEM)
        bm = 128 if tile else 64
        tiles = ((it["m"] + bm - 1) // bm) * ((it["n"] + bm - 1) // bm)
        self._blocks_ok(it, tiles, total, "gemm mixed")
        for r in it:
            m, nn, k = int(r["m"]), int(r["n"]), int(r["k"])
            A = self.mat(r["A"], m, k, int(r["lda"]), int(r["transA"]), "gemm A").astype(np.float32)
            Bm = self.mat(r["B"], k, nn, int(r["ldb"]), int(r["transB"]), "gemm B").astype(np.float32)
            C = self.mat(r["C"], m, nn, int(r["ldc"]), 0, "gemm C")
            res = float(r["alpha"]) * (A @ Bm).astype(np.float64)
            if float(r["beta"]) != 0.0:
                res = res + float(r["beta"]) * C
            C[...] = res
        return 0

    def syn_mpo_apply_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.APPLY_ITEM)
        nel = it["chil"] * it["dl"] * it["dout"] * it["chir"] * it["dr"]
        self._blocks_ok(it, nel, total, "apply", per=_lib.APPLY_BLOCK)
        for r in it:
            a, s, b, c, o, e = (int(r[f]) for f in ("chil", "din", "chir", "dl", "dout", "dr"))
            X = self.buf(r["x"], a * s * b, what="apply x").reshape(a, s, b)
            O = self.buf(r["o"], c * s * o * e, what="apply o").reshape(c, s, o, e)
            P = self.buf(r["p"], a * c * o * b * e, what="apply p")
            P[:] = np.einsum("asb,csoe->acobe", X, O).reshape(-1)
        return 0

    def syn_mpo_mpo_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.MPO_MPO_ITEM)
        nel = it["al"] * it["bl"] * it["din"] * it["dout"] * it["ar"] * it["br"]
        self._blocks_ok(it, nel, total, "mpo_mpo", per=_lib.APPLY_BLOCK)
        for r in it:
            al, di, dm, ar, bl, do, br = (int(r[f]) for f in ("al", "din", "dmid", "ar", "bl", "dout", "br"))
            A = self.buf(r["a"], al * di * dm * ar, what="mpo_mpo a").reshape(al, di, dm, ar)
            B = self.buf(r["b"], bl * dm * do * br, what="mpo_mpo b").reshape(bl, dm, do, br)
            P = self.buf(r["p"], al * bl * di * do * ar * br, what="mpo_mpo p")
            P[:] = np.einsum("aijb,cjpe->acipbe", A, B).reshape(-1)
        return 0

    def syn_eye_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.EYE_ITEM)
        self._blocks_ok(it, it["n"] ** 2, total, "eye", per=_lib.APPLY_BLOCK)
        for r in it:
            X = self.mat(r["x"], int(r["n"]), int(r["n"]), int(r["ld"]), 0, "eye")
            X[...] = np.eye(int(r["n"]))
        return 0

    def syn_copy_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.COPY_ITEM)
        self._blocks_ok(it, it["n"], total, "copy", per=_lib.APPLY_BLOCK)
        for r in it:
            src = self.buf(r["src"], int(r["n"]), what="copy src").copy()
            self.buf(r["dst"], int(r["n"]), what="copy dst")[:] = src
        return 0

    def syn_rowscale_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.ROWSCALE_ITEM)
        self._blocks_ok(it, it["rows"] * it["cols"], total, "rowscale", per=_lib.APPLY_BLOCK)
        for r in it:
            rows = int(r["rows"])
            X = self.mat(r["x"], rows, int(r["cols"]), int(r["ld"]), 0, "rowscale x")
            s = self.buf(r["s"], rows, what="rowscale s")
            if int(r["mode"]) == 0:
                X *= s[:, None]
            else:
                inv = np.where(s != 0, 1.0 / np.where(s != 0, s, 1.0), 0.0)
                X *= inv[:, None]
        return 0

    def syn_vecnorms_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.VECNORM_ITEM)
        blocks = np.where(it["axis"] == 0, (it["cols"] + 63) // 64, (it["rows"] + 3) // 4)
        st = np.concatenate([[0], np.cumsum(blocks)[:-1]])
        if not np.array_equal(st, it["block_start"]) or int(blocks.sum()) != int(total):
            raise BoundsError(f"vecnorms: block table {list(it['block_start'])} / {total} does not fit the items")
        for r in it:
            rows, cols = int(r["rows"]), int(r["cols"])
            X = self.mat(r["x"], rows, cols, int(r["ld"]), 0, "vecnorms x")
            axis = int(r["axis"])
            out = self.buf(r["out"], cols if axis == 0 else rows, what="vecnorms out")
            out[:] = np.sqrt((X * X).sum(axis=axis))
        return 0

    def syn_qr_f64(self, items, n, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.QR_ITEM)
        for r in it:
            m, nn, keep = int(r["m"]), int(r["n"]), int(r["keep"])
            nb, kq = min(nn, keep), min(keep, m)
            ext = (m - 1) * int(r["rs"]) + (nn - 1) * int(r["cs"]) + 1
            b = self.buf(r["a"], ext, what="qr a")
            A = np.lib.stride_tricks.as_strided(b, (m, nn), (8 * int(r["rs"]), 8 * int(r["cs"])))
            self.buf(r["work"], m * (nb + (kq if r["q_trans"] else 0)), what="qr work")
            self.buf(r["tau"], nb, what="qr tau")
            Q = np.linalg.qr(A[:, :nb], mode="complete")[0][:, :kq]
            if r["q_trans"]:
                self.mat(r["q"], kq, m, int(r["ldq"]), 0, "qr q")[...] = Q.T
            else:
                self.mat(r["q"], m, kq, int(r["ldq"]), 0, "qr q")[...] = Q
        return 0

    def syn_qr_sized_f64(self, items, n, lds_doubles, stream):
        if not 0 <= lds_doubles <= 12288:
            raise BoundsError("qr LDS size beyond 96 KB")
        return self.syn_qr_f64(items, n, stream)

    def _jacobi_one(self, x, ldx, nv, ln, rt, ldr, s, gram):
        X = self.mat(x, nv, ln, ldx, 0, "jacobi x")
        R = self.mat(rt, nv, nv, ldr, 0, "jacobi rt")
        S = self.buf(s, nv, what="jacobi s")
        if gram:
            if nv != ln:
                raise BoundsError("gram mode needs a square matrix")
            w, V = np.linalg.eigh(0.5 * (X + X.T))
            order = np.argsort(-w, kind="stable")
            w, V = w[order].clip(min=0), V[:, order]
            R[...] = V.T
            X[...] = V.T @ X
            S[:] = np.sqrt(w)
        else:
            U, sv, Vt = np.linalg.svd(X, full_matrices=False)
            R[...] = np.eye(nv)
            R[:len(sv)] = U.T
            X[...] = 0
            X[:len(sv)] = sv[:, None] * Vt
            S[:] = 0
            S[:len(sv)] = sv

    def syn_jacobi_f64(self, items, n, tol, max_sweeps, stream):
        from syngular_amd import _lib
        for r in self.table(items, n, _lib.JACOBI_ITEM):
            self._jacobi_one(r["x"], int(r["ldx"]), int(r["nv"]), int(r["len"]), r["rt"], int(r["ldr"]), r["s"],
                             int(r["gram"]))
            if int(r["sweeps"]):
                self.buf(r["sweeps"], 1, np.int32, "jacobi sweeps")[0] = 1
        return 0

    def syn_jacobi_gram_f64(self, items, n, nv_max, tol, max_sweeps, stream):
        from syngular_amd import _lib
        for r in self.table(items, n, _lib.JACOBI_ITEM):
            nv = int(r["nv"])
            if nv > nv_max or nv > 64 or int(r["len"]) != nv:
                raise BoundsError(f"gram jacobi item of size {nv} (nv_max {nv_max})")
            X = self.mat(r["x"], nv, nv, int(r["ldx"]), 0, "gram x")
            R = self.mat(r["rt"], nv, nv, int(r["ldr"]), 0, "gram rt")
            S = self.buf(r["s"], nv, what="gram s")
            w, V = np.linalg.eigh(0.5 * (X + X.T))
            order = np.argsort(-w, kind="stable")
            R[...] = V[:, order].T
            S[:] = np.sqrt(w[order].clip(min=0))
        return 0

    def syn_jacobi_blocked_many_f64(self, items_dev, items_host, n, work, counters, tol, max_sweeps, sweeps, stream):
        from syngular_amd import _lib
        self.buf(items_dev, n * _lib.BJAC_ITEM.itemsize, np.uint8, "bjac device table")
        addr = items_host.value if isinstance(items_host, ctypes.c_void_p) else int(items_host)
        host = (ctypes.c_char * (n * _lib.BJAC_ITEM.itemsize)).from_address(addr)
        it = np.frombuffer(bytes(host), dtype=_lib.BJAC_ITEM)
        nv = it["nv"].astype(np.int64)
        ln = it["len"].astype(np.int64)
        self.buf(work, int(_workspace(nv, ln)), what="bjac work")
        self.buf(counters, n, np.int32, "bjac counters")
        for r in it:
            self._jacobi_one(r["x"], int(r["ldx"]), int(r["nv"]), int(r["len"]), r["rt"], int(r["ldr"]), r["s"],
                             int(r["gram"]))
        if sweeps is not None and hasattr(sweeps, "_obj"):
            sweeps._obj.value = 1
        return 0

    def syn_diagmax_f64(self, items, n, stream):
        from syngular_amd import _lib
        for r in self.table(items, n, _lib.DIAGMAX_ITEM):
            nn, ld = int(r["n"]), int(r["ld"])
            G = self.mat(r["g"], nn, nn, ld, 0, "diagmax g")
            self.buf(r["out"], 1, what="diagmax out")[0] = float(np.max(np.diag(G)))
            self.buf(r["dropped"], 1, np.int32, "diagmax dropped")[0] = 0
        return 0

    def syn_chol_leaf_f64(self, items, n, stream):
        """Right-looking Cholesky with the pivot-drop rule, then inv(L~)."""
        from syngular_amd import _lib
        for r in self.table(items, n, _lib.CHOL_ITEM):
            nn = int(r["n"])
            if not 1 <= nn <= 64:
                raise BoundsError(f"chol leaf of size {nn}")
            G = self.mat(r["g"], nn, nn, int(r["ldg"]), 0, "chol g")
            Lo = self.mat(r["l"], nn, nn, int(r["ldl"]), 0, "chol l")
            Xo = self.mat(r["x"], nn, nn, int(r["ldx"]), 0, "chol x")
            thr = float(r["tolrel"]) * float(self.buf(r["maxdiag"], 1, what="chol maxdiag")[0])
            a = np.tril(G) + np.tril(G, -1).T
            L = np.zeros((nn, nn))
            keep = np.ones(nn, bool)
            for j in range(nn):
                d = a[j, j]
                if d > thr:
                    sd = np.sqrt(d)
                    L[j, j] = sd
                    L[j + 1:, j] = a[j + 1:, j] / sd
                    a[j + 1:, j + 1:] -= np.outer(a[j + 1:, j], a[j + 1:, j]) / d
                else:
                    keep[j] = False
            Lt = L.copy()
            for i in np.nonzero(~keep)[0]:
                Lt[i, :] = 0.0
                Lt[i, i] = 1.0
            Z = np.linalg.inv(Lt)
            Z[:, ~keep] = 0.0
            Lo[...] = L
            Xo[...] = np.tril(Z)
            self.buf(r["dropped"], 1, np.int32, "chol dropped")[0] += int((~keep).sum())
        return 0

    def _chol_drop(self, G, tolrel):
        """Cholesky with the pivot-drop rule; returns (L, X = inv(L~), drops)."""
        nn = G.shape[0]
        thr = tolrel * float(np.max(np.diag(G))) if nn else 0.0
        a = G.copy()
        L = np.zeros((nn, nn))
        keep = np.ones(nn, bool)
        for j in range(nn):
            d = a[j, j]
            if d > thr:
                sd = np.sqrt(d)
                L[j, j] = sd
                L[j + 1:, j] = a[j + 1:, j] / sd
                a[j + 1:, j + 1:] -= np.outer(a[j + 1:, j], a[j + 1:, j]) / d
            else:
                keep[j] = False
        Lt = L.copy()
        for i in np.nonzero(~keep)[0]:
            Lt[i, :] = 0.0
            Lt[i, i] = 1.0
        X = np.linalg.inv(Lt)
        X[:, ~keep] = 0.0
        return L, X, int((~keep).sum())

    def syn_cholqr_f64(self, a, m, n, lda, q, ldq, l, ldl, work, tolrel, dropped, mixed, stream):
        A = self.mat(a, m, n, lda, 0, "cholqr a").copy()
        L, X, drops = self._chol_drop(A.T @ A, tolrel)
        self.mat(q, m, n, ldq, 0, "cholqr q")[...] = A @ X.T
        self.mat(l, n, n, ldl, 0, "cholqr l")[...] = L
        self.buf(dropped, 1, np.int32, "cholqr dropped")[0] += drops
        return 0

    def syn_cholqr_refine_f64(self, q1, m, n, ldq1, q, ldq, l, ldl, work, thresh, tolrel, dropped, full, mixed,
                              stream):
        # the first-order branch of the device is an optimisation of the same
        # factorisation; the emulation always takes the Cholesky
        return self.syn_cholqr_f64(q1, m, n, ldq1, q, ldq, l, ldl, work, tolrel, dropped, mixed, stream)

    def syn_hh_panel_f64(self, w, ldw, v, ldv, tau, t, ldt, m, j0, b, stream):
        """Unblocked Householder on panel columns [j0, j0+b) (LAPACK
        convention), explicit V and the compact-WY T (dlarft, forward)."""
        if not (0 <= j0 < m and 1 <= b <= 32 and ldw >= j0 + b and ldv >= j0 + b and ldt >= b):
            raise BoundsError("hh panel arguments")
        W = self.mat(w, m, ldw, ldw, 0, "hh panel w")
        Vm = self.mat(v, m, ldv, ldv, 0, "hh panel v")
        taus = self.buf(tau, j0 + b, what="hh panel tau")
        T = self.mat(t, b, b, ldt, 0, "hh panel t")
        for jj in range(b):
            j = j0 + jj
            x = W[j + 1:, j]
            xn2 = float(x @ x)
            alpha = W[j, j]
            tj = 0.0
            if xn2 > 0.0:
                beta = -np.copysign(np.sqrt(alpha * alpha + xn2), alpha)
                tj = (beta - alpha) / beta
                W[j + 1:, j] = x / (alpha - beta)
                W[j, j] = beta
            taus[j] = tj
            col = np.zeros(m - j0)
            col[jj] = 1.0
            if tj != 0.0:
                col[jj + 1:] = W[j + 1:, j]
            Vm[j0:, j] = col
            if tj != 0.0 and jj + 1 < b:
                vv = col[jj:]
                C = W[j:, j + 1:j0 + b]
                C -= tj * np.outer(vv, vv @ C)
        Vp = Vm[j0:, j0:j0 + b]
        Tl = np.zeros((b, b))
        for jj in range(b):
            tj = taus[j0 + jj]
            Tl[jj, jj] = tj
            if jj:
                Tl[:jj, jj] = -tj * Tl[:jj, :jj] @ (Vp[:, :jj].T @ Vp[:, jj])
        T[...] = Tl
        return 0

    def syn_lu_sign_leaf_f64(self, items, n, amb, stream):
        from syngular_amd import _lib
        for r in self.table(items, n, _lib.LUSIGN_ITEM):
            bs = int(r["bs"])
            if not 1 <= bs <= 64:
                raise BoundsError(f"lu sign leaf of size {bs}")
            M = self.mat(r["a"], bs, bs, int(r["lda"]), 0, "lusign a").copy()
            D = self.buf(r["d"], bs, what="lusign d")
            amb_hit = False
            for j in range(bs):
                c = M[j, j]
                if j == int(r["free_col"]):
                    d = 1.0 if c >= 0 else -1.0
                else:
                    d = -1.0 if c >= 0 else 1.0
                    amb_hit |= abs(c) > amb
                D[j] = d
                M[j, j] = c - d
                M[j + 1:, j] /= M[j, j]
                M[j + 1:, j + 1:] -= np.outer(M[j + 1:, j], M[j, j + 1:])
            if int(r["linv"]):
                Lm = np.tril(M, -1) + np.eye(bs)
                self.mat(r["linv"], bs, bs, int(r["ldl"]), 0, "lusign linv")[...] = np.linalg.inv(Lm)
                self.mat(r["uinv"], bs, bs, int(r["ldu"]), 0, "lusign uinv")[...] = np.linalg.inv(np.triu(M))
            if amb_hit:
                self.buf(r["flag"], 1, np.int32, "lusign flag")[0] = 1
        return 0

    def syn_colscale_f64(self, items, n, total, stream):
        from syngular_amd import _lib
        it = self.table(items, n, _lib.COLSCALE_ITEM)
        self._blocks_ok(it, [((int(r["rows"]) + 63) // 64) * ((int(r["cols"]) + 63) // 64) for r in it], total,
                        "colscale")
        for r in it:
            rows, cols = int(r["rows"]), int(r["cols"])
            S = self.mat(r["src"], rows, cols, int(r["lds"]), 0, "colscale src").copy()
            if int(r["s"]):
                S = S * self.buf(r["s"], cols, what="colscale s")
            if int(r["trans"]):
                self.mat(r["dst"], cols, rows, int(r["ldd"]), 0, "colscale dst")[...] = S.T
            else:
                self.mat(r["dst"], rows, cols, int(r["ldd"]), 0, "colscale dst")[...] = S
        return 0

    def syn_jacobi_gram_blocked_f64(self, items_dev, items_host, n, work, counters, tol, max_sweeps, sweeps, stream):
        from syngular_amd import _lib
        self.buf(items_dev, n * _lib.GBJ_ITEM.itemsize, np.uint8, "gram device table")
        addr = items_host.value if isinstance(items_host, ctypes.c_void_p) else int(items_host)
        it = np.frombuffer(bytes((ctypes.c_char * (n * _lib.GBJ_ITEM.itemsize)).from_address(addr)),
                           dtype=_lib.GBJ_ITEM)
        self.buf(work, int(_gram_workspace(it["nv"])), what="gram work")
        self.buf(counters, n, np.int32, "gram counters")
        for r in it:
            nv = int(r["nv"])
            G = self.mat(r["g"], nv, nv, int(r["ldg"]), 0, "gram g")
            R = self.mat(r["rt"], nv, nv, int(r["ldr"]), 0, "gram rt")
            S = self.buf(r["s"], nv, what="gram s")
            w, V = np.linalg.eigh(0.5 * (G + G.T))
            order = np.argsort(-w, kind="stable")
            R[...] = V[:, order].T
            S[:] = np.sqrt(w[order].clip(min=0))
            G[...] = np.diag(w)
        if sweeps is not None and hasattr(sweeps, "_obj"):
            sweeps._obj.value = 1
        return 0

    def syn_truncate_f64(self, items, n, stream):
        from syngular_amd import _lib
        for r in self.table(items, n, _lib.TRUNC_ITEM):
            s = self.buf(r["s"], int(r["n"]), what="trunc s")
            c = keep_count(s.copy(), int(r["chi_max"]) or None, float(r["cutoff"]))
            self.buf(r["keep"], 1, np.int32, "trunc keep")[0] = c
            self.buf(r["discarded"], 1, what="trunc disc")[0] = float(np.sum(s[c:] ** 2))
        return 0


def _workspace(nv, ln):
    nv, ln = np.asarray(nv, np.int64), np.asarray(ln, np.int64)
    return int((nv * ln + nv * nv + nv + 2 * nv + 2).sum() + 2 * len(nv) + 16)


def _gram_workspace(nv):
    tot = 0
    for v in np.asarray(nv, np.int64):
        v = int(v)
        work, owners = [], []
        for tb in (16, 32):
            nblk = (v + tb - 1) // tb
            half = (nblk + (nblk & 1)) // 2
            # two W-tile sets, rotation flags, floor, sort scratch, second Gram buffer
            work.append(2 * half * 4 * tb * tb + half + 1 + v * v + v + 4 + v * v)
            gtiles = half * (half + 1) // 2
            cols = 128 if tb == 16 else 64        # rt slab width of the block size
            ugroups = ((gtiles + 1) // 2 if tb == 16 else gtiles) + half * ((v + cols - 1) // cols)
            owners.append(half + ugroups)
        tot += max(work) + 2 * v + (max(owners) + 1) // 2
    return tot + 4 * len(nv) + 16


class _FakeLib:
    def syn_cholqr_workspace(self, n):
        return 1

    def syn_jacobi_gram_blocked_workspace(self, nv_ptr, n):
        nv = np.ctypeslib.as_array(ctypes.cast(nv_ptr, ctypes.POINTER(ctypes.c_int64)), (n,))
        return _gram_workspace(nv)

    def syn_jacobi_blocked_many_workspace(self, nv_ptr, len_ptr, n):
        nv = np.ctypeslib.as_array(ctypes.cast(nv_ptr, ctypes.POINTER(ctypes.c_int64)), (n,))
        ln = np.ctypeslib.as_array(ctypes.cast(len_ptr, ctypes.POINTER(ctypes.c_int64)), (n,))
        return _workspace(nv, ln)


def install(monkeypatch, tensor_api=False):
    """Route syngular_amd.batch / serving through the emulator on the CPU.
    tensor_api: also route the single-problem calls of syngular_amd.ops (the
    MatrixProductState / Operator path).  Those allocate with torch directly,
    so accesses are not bounds-checked in that mode (the GPU tests cover the
    same calls)."""
    from syngular_amd import _lib, batch, ops

    emu = Emulator()
    if tensor_api:
        emu.checked = False
        monkeypatch.setattr(ops, "call", emu.call)
    cpu = torch.device("cpu")

    def empty(n, dtype=torch.float64, zero=False):
        f = torch.zeros if zero else torch.empty
        return emu.register(f(max(int(n), 1), dtype=dtype, device=cpu))

    def h2d(host):
        return emu.register(host.clone())

    def upload(items):
        raw = np.ascontiguousarray(items).view(np.uint8).reshape(-1)
        return emu.register(torch.from_numpy(raw.copy()))

    monkeypatch.setattr(batch, "_empty", empty)
    monkeypatch.setattr(batch, "_h2d", h2d)
    monkeypatch.setattr(batch, "call", emu.call)
    monkeypatch.setattr(ops, "_upload", upload)
    monkeypatch.setattr(ops, "device", lambda: cpu)
    monkeypatch.setattr(ops, "_stream", lambda: ctypes.c_void_p(0))
    monkeypatch.setattr(_lib, "load", lambda: _FakeLib())
    return emu


__all__ = ["Emulator", "install", "BoundsError"]
sys.modules.setdefault("abi_emulator", sys.modules[__name__])
