"""Device operations: thin Python wrappers that pack item tables and call the C ABI.

Every function works on float64 CUDA (HIP) tensors, launches on torch's
current stream and returns device tensors.  Item tables are packed on the host
as numpy structured arrays and copied to the device once per launch.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call

F64 = torch.float64
EPS = float(np.finfo(np.float64).eps)


# set once a device has been found and the library loaded: the checks then
# cost nothing on the per-launch path (a served layer makes thousands of
# launches from host threads that share one interpreter lock)
_gpu_ready = [False]


def require_gpu():
    if _gpu_ready[0]:
        return
    if not torch.cuda.is_available():
        raise _lib.SyngularError("no HIP device visible: syngular_amd runs on an MI355X (gfx950) only")
    _lib.load()
    _gpu_ready[0] = True


def device():
    require_gpu()
    return torch.device("cuda", torch._C._cuda_getDevice())


def _stream():
    """The calling thread's current HIP stream of the current device, as the
    raw handle the library's entry points take."""
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))


class StagingPool:
    """Pinned host memory for the item tables of a HIP graph capture.  No
    pinned allocation may happen while a stream captures, and every replay
    copies from the staging bytes again, so the tables are bump-allocated
    from one buffer allocated beforehand and kept for the graph's lifetime
    (``count_uploads`` measures the bytes in an eager run of the same work)."""

    ALIGN = 256

    def __init__(self, nbytes):
        self.buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, pin_memory=True)
        self.off = 0
        self.tables = 0

    def take(self, n):
        o = (self.off + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        if o + n > self.buf.numel():
            raise RuntimeError(f"staging pool of {self.buf.numel()} bytes exhausted (needs {o + n})")
        self.off = o + n
        self.tables += 1
        return self.buf[o:o + n]


# the StagingPool of the capture in progress (None when not capturing)
_staging = [None]
# bytes the uploads of an eager run would take from a pool (None: not counting)
_upload_count = [None]


def count_uploads(on=True):
    """Start (or stop and return) counting the pool bytes of the uploads."""
    if on:
        _upload_count[0] = 0
        return 0
    n, _upload_count[0] = _upload_count[0], None
    return n


def _upload(items: np.ndarray) -> torch.Tensor:
    """Copy an item table to the device without blocking the host: the bytes
    go through torch's cached pinned-host allocator, which keeps the staging
    buffer alive until the asynchronous copy on the current stream is done;
    under graph capture they go through the capture's StagingPool."""
    raw = np.ascontiguousarray(items).view(np.uint8).reshape(-1)
    if _upload_count[0] is not None:
        _upload_count[0] += (raw.size + StagingPool.ALIGN - 1) // StagingPool.ALIGN * StagingPool.ALIGN
    pool = _staging[0]
    if pool is not None:
        host = pool.take(raw.size)
        host.numpy()[:] = raw
        dev = torch.empty(raw.size, dtype=torch.uint8, device=device())
        dev.copy_(host, non_blocking=True)
        return dev
    host = torch.empty(raw.size, dtype=torch.uint8, pin_memory=True)
    host.numpy()[:] = raw
    return host.to(device(), non_blocking=True)


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def as_matrix(t: torch.Tensor):
    """(ptr, ld, trans) so that element (i, j) of the 2-D view t is
    ptr[i*ld + j] (trans=0) or ptr[j*ld + i] (trans=1)."""
    if t.dim() != 2:
        raise ValueError("expected a matrix")
    if t.dtype != F64:
        raise TypeError(f"expected float64, got {t.dtype}")
    r, c = t.shape
    s0, s1 = t.stride()
    if s1 == 1 or c == 1:
        return _ptr(t), (s0 if r > 1 else max(c, 1)), 0
    if s0 == 1 or r == 1:
        return _ptr(t), (s1 if c > 1 else max(r, 1)), 1
    raise ValueError("matrix view must be row- or column-contiguous")


# --------------------------------------------------------------------------
# GEMM
# --------------------------------------------------------------------------
_TILES = {0: 64, 1: 128}


def gemm_fn(mixed):
    """Entry point of a compute kind: False (FP64), True (FP32 matrix cores)
    or "bf16" (BF16 matrix cores); FP64 operands and results in every case."""
    if isinstance(mixed, str):
        if mixed != "bf16":
            raise ValueError(f"unknown compute kind {mixed!r}")
        return "syn_gemm_bf16_f64"
    return "syn_gemm_mixed_f64" if mixed else "syn_gemm_f64"


def gemm_many(problems, tile=None, mixed=False):
    """problems: iterable of (A, B, C, alpha, beta[, tri]) with A (m x k),
    B (k x n), C (m x n) 2-D float64 views; C is overwritten with
    alpha*A@B + beta*C.  tri is the structural-zero mask of syn_gemm_item.
    mixed: products on the FP32 matrix cores (operands rounded to FP32, FP32
    accumulation, FP64 result), or "bf16" for the BF16 matrix cores."""
    problems = [p if len(p) == 6 else (*p, 0) for p in problems if p[2].numel() > 0]
    if not problems:
        return
    if tile is None:
        work = sum(p[2].shape[0] * p[2].shape[1] for p in problems)
        tile = 1 if work >= 256 * 128 * 128 else 0
    bm = _TILES[tile]
    items = np.zeros(len(problems), dtype=_lib.GEMM_ITEM)
    start = 0
    for i, (A, B, C, alpha, beta, tri) in enumerate(problems):
        m, k = A.shape
        k2, n = B.shape
        if k != k2 or C.shape != (m, n):
            raise ValueError(f"gemm shape mismatch {tuple(A.shape)} @ {tuple(B.shape)} -> {tuple(C.shape)}")
        pa, lda, ta = as_matrix(A)
        pb, ldb, tb = as_matrix(B)
        pc, ldc, tc = as_matrix(C)
        if tc:
            raise ValueError("gemm output must be row-major")
        tn = (n + bm - 1) // bm
        tm = (m + bm - 1) // bm
        it = items[i]
        it["A"], it["B"], it["C"] = pa, pb, pc
        it["m"], it["n"], it["k"] = m, n, k
        it["lda"], it["ldb"], it["ldc"] = lda, ldb, ldc
        it["alpha"], it["beta"] = alpha, beta
        it["transA"], it["transB"] = ta, tb
        it["block_start"], it["tiles_n"] = start, tn
        it["tri"] = tri
        start += tm * tn
    dev = _upload(items)
    if _profiling[0]:
        frac = {0: 1.0, 1: 0.5}
        fl = sum(2.0 * p[0].shape[0] * p[0].shape[1] * p[1].shape[1] * frac.get(bin(p[5]).count("1"), 1.0 / 3.0)
                 for p in problems)
        call("syn_prof_note_flops", fl)
    call(gemm_fn(mixed), _ptr(dev), len(problems), start, tile, _stream())
    return dev


_profiling = [False]


def profile_begin(tags=None):
    """Start live per-kernel event timing inside the library, of the kernels
    named in `tags` (all instrumented kernels when None)."""
    _profiling[0] = True
    _lib.prof_begin(tags)


def profile_end():
    """Stop timing; {kernel: (launches, ms, algorithmic flops)}."""
    _profiling[0] = False
    return _lib.prof_end()


def gemm(A, B, C=None, alpha=1.0, beta=0.0, tile=None, tri=0, mixed=False):
    if C is None:
        C = torch.empty((A.shape[0], B.shape[1]), dtype=F64, device=A.device)
        beta = 0.0
    gemm_many([(A, B, C, alpha, beta, tri)], tile=tile, mixed=mixed)
    return C


# --------------------------------------------------------------------------
# site contractions and packing
# --------------------------------------------------------------------------
def _blocks(nel):
    return (nel + _lib.APPLY_BLOCK - 1) // _lib.APPLY_BLOCK


def apply_many(pairs):
    """pairs: list of (O (dl,din,dout,dr), X (chil,din,chir)); returns the
    materialised cores P (chil*dl, dout, chir*dr)."""
    outs = []
    items = np.zeros(len(pairs), dtype=_lib.APPLY_ITEM)
    start = 0
    for i, (O, X) in enumerate(pairs):
        dl, din, dout, dr = O.shape
        chil, din2, chir = X.shape
        if din != din2:
            raise ValueError(f"physical dimension mismatch: MPO input {din} vs MPS {din2}")
        O = O.contiguous()
        X = X.contiguous()
        P = torch.empty((chil * dl, dout, chir * dr), dtype=F64, device=X.device)
        outs.append((P, O, X))
        it = items[i]
        it["x"], it["o"], it["p"] = _ptr(X), _ptr(O), _ptr(P)
        it["chil"], it["din"], it["chir"], it["dl"], it["dout"], it["dr"] = chil, din, chir, dl, dout, dr
        it["block_start"] = start
        start += _blocks(P.numel())
    if start:
        dev = _upload(items)
        if _profiling[0]:
            # bytes moved: every output element written once, operands read once
            byt = sum(8.0 * (o[0].numel() + o[1].numel() + o[2].numel()) for o in outs)
            call("syn_prof_note_flops", byt)
        call("syn_mpo_apply_f64", _ptr(dev), len(pairs), start, _stream())
    return [o[0] for o in outs]


def mpo_mpo_many(pairs):
    """pairs: list of (A (al,din,dmid,ar), B (bl,dmid,dout,br)); returns
    P (al*bl, din, dout, ar*br) with P[a*bl+c, i, p, b*br+e] = sum_j A[a,i,j,b] B[c,j,p,e]."""
    outs = []
    items = np.zeros(len(pairs), dtype=_lib.MPO_MPO_ITEM)
    start = 0
    for i, (A, B) in enumerate(pairs):
        al, din, dm, ar = A.shape
        bl, dm2, dout, br = B.shape
        if dm != dm2:
            raise ValueError(f"operator dimension mismatch {dm} vs {dm2}")
        A = A.contiguous()
        B = B.contiguous()
        P = torch.empty((al * bl, din, dout, ar * br), dtype=F64, device=A.device)
        outs.append((P, A, B))
        it = items[i]
        it["a"], it["b"], it["p"] = _ptr(A), _ptr(B), _ptr(P)
        it["al"], it["din"], it["dmid"], it["ar"] = al, din, dm, ar
        it["bl"], it["dout"], it["br"] = bl, dout, br
        it["block_start"] = start
        start += _blocks(P.numel())
    if start:
        dev = _upload(items)
        call("syn_mpo_mpo_f64", _ptr(dev), len(pairs), start, _stream())
    return [o[0] for o in outs]


def kron_bonds(A, B):
    """P[(a*bl+c), i, o, (b*br+e)] = A[a,i,o,b] * B[c,i,o,e] (one launch)."""
    al, din, dout, ar = A.shape
    bl, _, _, br = B.shape
    A = A.contiguous()
    B = B.contiguous()
    P = torch.empty((al * bl, din, dout, ar * br), dtype=F64, device=A.device)
    items = np.zeros(1, dtype=_lib.MPO_MPO_ITEM)
    it = items[0]
    it["a"], it["b"], it["p"] = _ptr(A), _ptr(B), _ptr(P)
    it["al"], it["din"], it["dmid"], it["ar"] = al, din, dout, ar
    it["bl"], it["dout"], it["br"] = bl, dout, br
    nb = _blocks(P.numel())
    if nb:
        dev = _upload(items)
        call("syn_kron_bonds_f64", _ptr(dev), 1, nb, _stream())
    return P


def direct_sum_many(triples):
    """triples: list of (A, B, edge) with cores (l, *mid, r); edge 1 = first
    site, 2 = last site, 0 = interior.  Returns the packed cores."""
    outs = []
    items = np.zeros(len(triples), dtype=_lib.DSUM_ITEM)
    start = 0
    for i, (A, B, edge) in enumerate(triples):
        A = A.contiguous()
        B = B.contiguous()
        mid_shape = tuple(A.shape[1:-1])
        if tuple(B.shape[1:-1]) != mid_shape:
            raise ValueError("could not broadcast: physical dimensions differ")
        mid = int(np.prod(mid_shape))
        la, ra, lb, rb = A.shape[0], A.shape[-1], B.shape[0], B.shape[-1]
        if edge == 1:
            if la != lb:
                raise ValueError("could not broadcast: left boundary bonds differ")
            lo, ro = la, ra + rb
        elif edge == 2:
            if ra != rb:
                raise ValueError("could not broadcast: right boundary bonds differ")
            lo, ro = la + lb, ra
        else:
            lo, ro = la + lb, ra + rb
        out = torch.empty((lo, *mid_shape, ro), dtype=F64, device=A.device)
        outs.append((out, A, B))
        it = items[i]
        it["a"], it["b"], it["out"] = _ptr(A), _ptr(B), _ptr(out)
        it["la"], it["ra"], it["lb"], it["rb"], it["mid"], it["lo"], it["ro"] = la, ra, lb, rb, mid, lo, ro
        it["edge"] = edge
        it["block_start"] = start
        start += _blocks(out.numel())
    if start:
        dev = _upload(items)
        call("syn_direct_sum_f64", _ptr(dev), len(triples), start, _stream())
    return [o[0] for o in outs]


def rowscale_many(problems):
    """problems: list of (X (rows x cols view, row-major), s (rows,), mode)."""
    items = np.zeros(len(problems), dtype=_lib.ROWSCALE_ITEM)
    start = 0
    for i, (X, s, mode) in enumerate(problems):
        px, ld, tr = as_matrix(X)
        if tr:
            raise ValueError("rowscale needs a row-major view")
        it = items[i]
        it["x"], it["rows"], it["cols"], it["ld"] = px, X.shape[0], X.shape[1], ld
        it["s"], it["mode"], it["block_start"] = _ptr(s), mode, start
        start += _blocks(X.numel())
    if start:
        dev = _upload(items)
        call("syn_rowscale_f64", _ptr(dev), len(problems), start, _stream())


def vecnorms_many(problems):
    """problems: list of (X (rows x cols row-major view), axis, out): out
    (a contiguous float64 vector) gets the column norms of X (axis 0) or its
    row norms (axis 1)."""
    items = np.zeros(len(problems), dtype=_lib.VECNORM_ITEM)
    start = 0
    for i, (X, axis, out) in enumerate(problems):
        px, ld, tr = as_matrix(X)
        if tr:
            raise ValueError("vecnorms needs a row-major view")
        rows, cols = X.shape
        if axis not in (0, 1) or out.numel() < (cols if axis == 0 else rows) or not out.is_contiguous():
            raise ValueError(f"vecnorms: output of {out.numel()} for axis {axis} of a {rows} x {cols} matrix")
        it = items[i]
        it["x"], it["rows"], it["cols"], it["ld"] = px, rows, cols, ld
        it["axis"], it["out"], it["block_start"] = axis, _ptr(out), start
        start += (cols + 63) // 64 if axis == 0 else (rows + 3) // 4
    if start:
        dev = _upload(items)
        call("syn_vecnorms_f64", _ptr(dev), len(problems), start, _stream())


def copy_many(pairs):
    """pairs: list of (src, dst) contiguous device tensors of equal size."""
    items = np.zeros(len(pairs), dtype=_lib.COPY_ITEM)
    start = 0
    for i, (src, dst) in enumerate(pairs):
        if src.numel() != dst.numel() or not (src.is_contiguous() and dst.is_contiguous()):
            raise ValueError("copy needs contiguous tensors of equal size")
        it = items[i]
        it["src"], it["dst"], it["n"], it["block_start"] = _ptr(src), _ptr(dst), src.numel(), start
        start += _blocks(src.numel())
    if start:
        dev = _upload(items)
        call("syn_copy_f64", _ptr(dev), len(pairs), start, _stream())


def eye_many(views):
    """views: list of square row-major matrix views, each set to the identity."""
    items = np.zeros(len(views), dtype=_lib.EYE_ITEM)
    start = 0
    for i, X in enumerate(views):
        px, ld, tr = as_matrix(X)
        if tr or X.shape[0] != X.shape[1]:
            raise ValueError("eye needs a square row-major view")
        it = items[i]
        it["x"], it["n"], it["ld"], it["block_start"] = px, X.shape[0], ld, start
        start += _blocks(X.shape[0] * X.shape[0])
    if start:
        dev = _upload(items)
        call("syn_eye_f64", _ptr(dev), len(views), start, _stream())


# --------------------------------------------------------------------------
# factorizations
# --------------------------------------------------------------------------
def qr_many(problems, q_trans=False):
    """problems: list of (A (m x n view), keep).  Returns, per problem, the
    first kq = min(keep, m) columns of the complete Householder Q (m x kq),
    or its transpose (kq x m) when q_trans."""
    items = np.zeros(len(problems), dtype=_lib.QR_ITEM)
    outs = []
    for i, (A, keep) in enumerate(problems):
        m, n = A.shape
        if keep <= 0:
            raise ValueError("keep must be positive")
        if A.dtype != F64:
            raise TypeError("qr expects float64")
        nb = min(n, keep)
        kq = min(keep, m)
        Q = torch.empty((kq, m) if q_trans else (m, kq), dtype=F64, device=A.device)
        work = torch.empty(max(m * (nb + (kq if q_trans else 0)), 1), dtype=F64, device=A.device)
        tau = torch.empty(max(nb, 1), dtype=F64, device=A.device)
        outs.append((Q, work, tau, A))
        s0, s1 = A.stride()
        it = items[i]
        it["a"], it["rs"], it["cs"] = _ptr(A), s0, s1
        it["m"], it["n"], it["keep"] = m, n, keep
        it["q"], it["ldq"], it["q_trans"] = _ptr(Q), (m if q_trans else kq), int(q_trans)
        it["work"], it["tau"] = _ptr(work), _ptr(tau)
    if len(problems):
        dev = _upload(items)
        call("syn_qr_f64", _ptr(dev), len(problems), _stream())
    return [o[0] for o in outs]


def colscale_many(problems):
    """problems: list of (src (rows x cols row-major view), dst, s (cols,) or
    None, trans): dst = src diag(s), or its transpose (cols x rows) when
    trans; s None copies."""
    items = np.zeros(len(problems), dtype=_lib.COLSCALE_ITEM)
    start = 0
    for i, (src, dst, sc, trans) in enumerate(problems):
        ps, lds, ts = as_matrix(src)
        pd, ldd, td = as_matrix(dst)
        rows, cols = src.shape
        if ts or td:
            raise ValueError("colscale needs row-major views")
        if tuple(dst.shape) != ((cols, rows) if trans else (rows, cols)):
            raise ValueError(f"colscale: destination {tuple(dst.shape)} does not fit source {(rows, cols)}")
        it = items[i]
        it["src"], it["lds"], it["dst"], it["ldd"] = ps, lds, pd, ldd
        it["rows"], it["cols"], it["s"], it["trans"] = rows, cols, (0 if sc is None else _ptr(sc)), int(trans)
        it["block_start"] = start
        start += ((rows + 63) // 64) * ((cols + 63) // 64)
    if start:
        dev = _upload(items)
        call("syn_colscale_f64", _ptr(dev), len(problems), start, _stream())


def transpose(X):
    """Row-major copy of X^T (X a row- or column-major matrix view)."""
    px, ld, tr = as_matrix(X)
    if tr:
        # X is the transpose of a row-major (cols x rows) view: copy that
        X = X.t()
        out = torch.empty((X.shape[0], X.shape[1]), dtype=F64, device=X.device)
        colscale_many([(X, out, None, False)])
        return out
    out = torch.empty((X.shape[1], X.shape[0]), dtype=F64, device=X.device)
    colscale_many([(X, out, None, True)])
    return out


# Blocked Householder QR: panels of this many columns factored by
# syn_hh_panel_f64, everything else on the GEMM (compact WY form).
HH_PANEL = 32
# Problems with at least this many reflectors whose leading block does not
# fit the one-work-group kernel's LDS (96 KB) take the blocked path.
HH_BLOCKED_MIN_REFLECTORS = 64
HH_LDS_DOUBLES = 96 * 1024 // 8


def _row_major_copy(X):
    """Contiguous row-major copy of a matrix view (GEMM with the identity for
    strided or transposed views, a flat copy otherwise)."""
    if X.is_contiguous():
        out = torch.empty_like(X)
        copy_many([(X.reshape(-1), out.reshape(-1))])
        return out
    I = torch.empty((X.shape[1], X.shape[1]), dtype=F64, device=X.device)
    eye_many([I])
    return gemm(X, I)


def qr_blocked(A, keep, q_trans=False):
    """Same result as qr_many([(A, keep)], q_trans)[0] (the first
    kq = min(keep, m) columns of the complete Householder Q of A, LAPACK
    convention), by the blocked algorithm: p = min(m, n, keep) reflectors,
    panels of HH_PANEL columns factored by syn_hh_panel_f64, the columns
    right of each panel updated as C -= V (T^T (V^T C)), and Q accumulated
    backwards from [I; 0] as Q_sub -= V (T (V^T Q_sub)) (for Q^T:
    Q^T_sub -= ((Q^T_sub V) T^T) V^T), all products on the MFMA GEMM."""
    m, n = A.shape
    if keep <= 0:
        raise ValueError("keep must be positive")
    if A.dtype != F64:
        raise TypeError("qr expects float64")
    nb = min(n, keep)
    kq = min(keep, m)
    p = min(m, nb)
    dev = A.device
    W = _row_major_copy(A[:, :nb])
    V = torch.empty((m, nb), dtype=F64, device=dev)
    tau = torch.empty(max(nb, 1), dtype=F64, device=dev)
    panels = [(j0, min(HH_PANEL, p - j0)) for j0 in range(0, p, HH_PANEL)]
    T = torch.empty((len(panels), HH_PANEL, HH_PANEL), dtype=F64, device=dev)
    for q, (j0, b) in enumerate(panels):
        call("syn_hh_panel_f64", _ptr(W), nb, _ptr(V), nb, _ptr(tau), _ptr(T[q]), HH_PANEL, m, j0, b, _stream())
        if j0 + b < nb:
            Vp, C = V[j0:, j0:j0 + b], W[j0:, j0 + b:]
            Y = gemm(Vp.t(), C)
            Z = gemm(T[q][:b, :b].t(), Y)
            gemm(Vp, Z, C=C, alpha=-1.0, beta=1.0)
    if q_trans:
        Q = torch.zeros((kq, m), dtype=F64, device=dev)
        eye_many([Q[:, :kq]])
    else:
        Q = torch.zeros((m, kq), dtype=F64, device=dev)
        eye_many([Q[:kq, :]])
    for q in range(len(panels) - 1, -1, -1):
        j0, b = panels[q]
        if j0 >= kq:
            continue
        Vp, Tq = V[j0:, j0:j0 + b], T[q][:b, :b]
        if q_trans:
            Qs = Q[j0:, j0:]                               # (kq - j0) x (m - j0)
            Y = gemm(Qs, Vp)
            Z = gemm(Y, Tq.t())
            gemm(Z, Vp.t(), C=Qs, alpha=-1.0, beta=1.0)
        else:
            Qs = Q[j0:, j0:]                               # (m - j0) x (kq - j0)
            Y = gemm(Vp.t(), Qs)
            Z = gemm(Tq, Y)
            gemm(Vp, Z, C=Qs, alpha=-1.0, beta=1.0)
    return Q


# Householder QR through CholeskyQR2: from this many kept columns on, the Q
# factor comes from two CholeskyQR passes on the MFMA GEMM and the LAPACK signs
# from a blocked LU of its top block; fewer take the one-work-group kernel.
HH_CHOLQR_MIN = 64
# Switched off: the first device run of this path ended in a GPU memory fault
# whose cause is not identified (the kernels' and the host flow's indexing
# check in bounds on the CPU).  Until it is, every QR takes the one-work-group
# Householder kernel.
HH_FAST_PATH = False
# |c_j| above this marks a column that LAPACK may leave unreflected (its part
# below the diagonal exactly zero): the one-work-group kernel decides those.
HH_SIGN_AMBIGUOUS = 1.0 - 1e-10
LU_LEAF = 64


def lu_signs(Qtop, m, flag):
    """Householder signs D (nb,) for a positive-diagonal orthonormal factor
    whose top nb x nb block is Qtop (row-major view), of an m-row matrix: the
    blocked right-looking LU without pivoting of Qtop - D with D chosen per
    column (syn_lu_sign_leaf_f64 on 64-column diagonal blocks, the GEMM for
    the panels and the trailing update).  flag (1,) int32 is set when a sign
    is ambiguous."""
    nb = Qtop.shape[0]
    dev = Qtop.device
    M = torch.empty((nb, nb), dtype=F64, device=dev)
    colscale_many([(Qtop, M, None, False)])
    D = torch.empty(nb, dtype=F64, device=dev)
    inv = torch.empty((2, LU_LEAF, LU_LEAF), dtype=F64, device=dev)
    for j0 in range(0, nb, LU_LEAF):
        bs = min(LU_LEAF, nb - j0)
        last = j0 + bs >= nb
        it = np.zeros(1, dtype=_lib.LUSIGN_ITEM)
        it["a"], it["lda"], it["bs"] = _ptr(M) + 8 * (j0 * nb + j0), nb, bs
        it["free_col"] = (m - 1 - j0) if j0 <= m - 1 < j0 + bs else -1
        it["d"] = _ptr(D) + 8 * j0
        it["linv"], it["ldl"] = (0 if last else _ptr(inv[0])), LU_LEAF
        it["uinv"], it["ldu"] = _ptr(inv[1]), LU_LEAF
        it["flag"] = _ptr(flag)
        call("syn_lu_sign_leaf_f64", _ptr(_upload(it)), 1, HH_SIGN_AMBIGUOUS, _stream())
        if last:
            break
        e = j0 + bs
        L21 = gemm(M[e:, j0:e], inv[1][:bs, :bs])         # M21 U^-1
        U12 = gemm(inv[0][:bs, :bs], M[j0:e, e:])         # L^-1 M12
        gemm(L21, U12, C=M[e:, e:], alpha=-1.0, beta=1.0)
    return D


def householder_q(A, keep, q_trans=False):
    """The first kq = min(keep, m) columns of the complete Householder Q of A
    (m x n view) in the LAPACK convention, i.e. numpy.linalg.qr(A,
    'complete')[0][:, :keep]; its transpose (kq x m) when q_trans.  Same
    result as qr_many([(A, keep)], q_trans)[0].

    With nb = min(n, keep) >= HH_CHOLQR_MIN and keep <= min(m, n) the first
    keep columns of Q are those of the Householder QR of A[:, :keep], taken as
    CholeskyQR2 (MFMA GEMM) times the sign matrix from lu_signs.  A pivot
    dropped by the first pass (a rank-deficient block) or an ambiguous sign
    sends the problem to the one-work-group Householder kernel."""
    m, n = A.shape
    if keep <= 0:
        raise ValueError("keep must be positive")
    if A.dtype != F64:
        raise TypeError("qr expects float64")
    nb = min(n, keep)
    if not HH_FAST_PATH or nb < HH_CHOLQR_MIN or keep > n or keep > m:
        if min(m, nb) >= HH_BLOCKED_MIN_REFLECTORS and m * nb > HH_LDS_DOUBLES:
            return qr_blocked(A, keep, q_trans)
        return qr_many([(A, keep)], q_trans=q_trans)[0]
    Ak = A[:, :nb]
    if as_matrix(Ak)[2]:
        Ak = transpose(Ak.t())
    Q1, L1, dropped = cholqr(Ak)
    if int(dropped.item()):
        return qr_many([(A, keep)], q_trans=q_trans)[0]
    Q, _, _ = cholqr_refine(Q1)
    flag = torch.zeros(1, dtype=torch.int32, device=A.device)
    D = lu_signs(Q[:nb], m, flag)
    if int(flag.item()):
        return qr_many([(A, keep)], q_trans=q_trans)[0]
    out = torch.empty((nb, m) if q_trans else (m, nb), dtype=F64, device=A.device)
    colscale_many([(Q, out, D, q_trans)])
    return out


JACOBI_TOL = 1e-15
JACOBI_SWEEPS = 60


def jacobi_many(mats, tol=None, max_sweeps=JACOBI_SWEEPS, gram=False):
    """mats: list of row-major (nv x len) float64 tensors, orthogonalised in
    place.  Returns per problem (rt (nv x nv), s (nv,), sweeps (1,) int32) with
    x_in = rt^T x_out and rows of x_out sorted by decreasing norm s.  With
    gram=True each input is a Gram matrix C^T C and s holds the singular
    values of C (square roots of the row norms)."""
    if tol is None:
        tol = jacobi_tol(max((m.shape[1] for m in mats), default=1))
    items = np.zeros(len(mats), dtype=_lib.JACOBI_ITEM)
    outs = []
    for i, X in enumerate(mats):
        nv, ln = X.shape
        px, ldx, tr = as_matrix(X)
        if tr:
            raise ValueError("jacobi needs a row-major matrix")
        rt = torch.empty((nv, nv), dtype=F64, device=X.device)
        s = torch.empty(nv, dtype=F64, device=X.device)
        sw = torch.zeros(1, dtype=torch.int32, device=X.device)
        outs.append((rt, s, sw))
        it = items[i]
        it["x"], it["ldx"], it["nv"], it["len"] = px, ldx, nv, ln
        it["rt"], it["ldr"], it["s"], it["sweeps"] = _ptr(rt), nv, _ptr(s), _ptr(sw)
        it["gram"] = int(gram)
    if len(mats):
        dev = _upload(items)
        call("syn_jacobi_f64", _ptr(dev), len(mats), float(tol), int(max_sweeps), _stream())
    return outs


GRAM_LDS_MAX = 64


def jacobi_gram_many(mats, tol=None, max_sweeps=JACOBI_SWEEPS):
    """Two-sided Jacobi eigen-decomposition of symmetric Gram matrices
    (nv x nv, nv <= 64) in LDS.  Returns per problem (rt (nv x nv), s (nv,))
    with rt rows the eigenvectors by decreasing eigenvalue and s the square
    roots of the eigenvalues (the singular values of C for G = C^T C).  The
    inputs are not modified."""
    if tol is None:
        tol = jacobi_tol(GRAM_LDS_MAX)
    items = np.zeros(len(mats), dtype=_lib.JACOBI_ITEM)
    outs = []
    nv_max = 1
    for i, G in enumerate(mats):
        nv, ln = G.shape
        if nv != ln or nv > GRAM_LDS_MAX:
            raise ValueError(f"jacobi_gram_many takes square matrices up to {GRAM_LDS_MAX}, got {tuple(G.shape)}")
        px, ldx, tr = as_matrix(G)
        if tr:
            raise ValueError("jacobi_gram_many needs a row-major matrix")
        rt = torch.empty((nv, nv), dtype=F64, device=G.device)
        s = torch.empty(nv, dtype=F64, device=G.device)
        outs.append((rt, s))
        nv_max = max(nv_max, nv)
        it = items[i]
        it["x"], it["ldx"], it["nv"], it["len"] = px, ldx, nv, nv
        it["rt"], it["ldr"], it["s"], it["sweeps"] = _ptr(rt), nv, _ptr(s), 0
        it["gram"] = 1
    if len(mats):
        dev = _upload(items)
        call("syn_jacobi_gram_f64", _ptr(dev), len(mats), nv_max, float(tol), int(max_sweeps), _stream())
    return outs


def truncate_many(problems):
    """problems: list of (s sorted desc, chi_max or None, cutoff).  Returns
    (keep int32 (n,), discarded f64 (n,)) device tensors."""
    n = len(problems)
    dev_ = device()
    keep = torch.zeros(n, dtype=torch.int32, device=dev_)
    disc = torch.zeros(n, dtype=F64, device=dev_)
    items = np.zeros(n, dtype=_lib.TRUNC_ITEM)
    for i, (s, chi_max, cutoff) in enumerate(problems):
        it = items[i]
        it["s"], it["n"] = _ptr(s), s.numel()
        it["chi_max"] = 0 if chi_max is None else int(chi_max)
        it["cutoff"] = float(cutoff or 0.0)
        it["keep"] = _ptr(keep) + 4 * i
        it["discarded"] = _ptr(disc) + 8 * i
    if n:
        dev = _upload(items)
        call("syn_truncate_f64", _ptr(dev), n, _stream())
    return keep, disc


# --------------------------------------------------------------------------
# large-problem paths (multi-step routines inside the library)
# --------------------------------------------------------------------------
def cholqr_tol(m, n):
    """Pivot-drop threshold relative to max_i ||a_i||^2: the rounding noise of
    a Gram pivot is a modest multiple of (m + n) eps."""
    return 8.0 * (m + n) * EPS


def cholqr(A, tolrel=None, mixed=False):
    """One CholeskyQR pass on a row-major m x n matrix (A is not modified).
    Returns (Q (m x n), L (n x n lower), dropped (1,) int32) with A = Q L^T.
    mixed: Q = A L^{-T} on the FP32 matrix cores (Gram and Cholesky in FP64)."""
    m, n = A.shape
    if tolrel is None:
        tolrel = cholqr_tol(m, n)
    pa, lda, tr = as_matrix(A)
    if tr:
        raise ValueError("cholqr needs a row-major matrix")
    Q = torch.empty((m, n), dtype=F64, device=A.device)
    L = torch.empty((n, n), dtype=F64, device=A.device)
    work = torch.empty(int(_lib.load().syn_cholqr_workspace(n)), dtype=F64, device=A.device)
    dropped = torch.zeros(1, dtype=torch.int32, device=A.device)
    call("syn_cholqr_f64", pa, m, n, lda, _ptr(Q), n, _ptr(L), n, _ptr(work), float(tolrel), _ptr(dropped),
         int(bool(mixed)), _stream())
    return Q, L, dropped


# Second CholeskyQR pass: when the first pass leaves max |Q1^T Q1 - I| below
# this, the Cholesky of I + E is taken to first order (error O(n 1e-18)).
CHOLQR_REFINE_THRESH = 1e-9
# Mixed form: its products round to FP32, so max|Q^T Q - I| after either
# pass is ~1e-7 whatever factor the second pass uses.  The first-order factor
# is exact to O(n max|E|^2), at most 4e-9 relative for n = 4096 and
# max|E| <= 1e-6: below the FP32 product error, so the full Cholesky of the
# second pass buys nothing there.
CHOLQR_REFINE_THRESH_MIXED = 1e-6


def cholqr_refine(Q1, tolrel=None, thresh=CHOLQR_REFINE_THRESH, mixed=False):
    """Second CholeskyQR pass on Q1 (not modified).  Returns (Q, L, full)
    with Q1 = Q L^T; full tells whether the Cholesky factorisation ran."""
    m, n = Q1.shape
    if tolrel is None:
        tolrel = cholqr_tol(m, n)
    pa, lda, tr = as_matrix(Q1)
    if tr:
        raise ValueError("cholqr_refine needs a row-major matrix")
    Q = torch.empty((m, n), dtype=F64, device=Q1.device)
    L = torch.empty((n, n), dtype=F64, device=Q1.device)
    work = torch.empty(int(_lib.load().syn_cholqr_workspace(n)), dtype=F64, device=Q1.device)
    dropped = torch.zeros(1, dtype=torch.int32, device=Q1.device)
    full = ctypes.c_int(0)
    call("syn_cholqr_refine_f64", pa, m, n, lda, _ptr(Q), n, _ptr(L), n, _ptr(work), float(thresh), float(tolrel),
         _ptr(dropped), ctypes.byref(full), int(bool(mixed)), _stream())
    return Q, L, bool(full.value)


# Adaptive second pass (cholqr2_adaptive): the first pass's loss of
# orthogonality E = Q1^T Q1 - I is estimated as ||Q1^T Q1 X - X||_F / sqrt(p)
# over p Gaussian probe columns X (an unbiased estimate of ||E||_F^2 in
# square; the chance that it falls below a tenth of ||E||_F is under 1e-7
# for p = 8), at the cost of two skinny products instead of a Gram matrix.
REORTH_PROBES = 8
_probe_cache = {}


def orthogonality_estimate(Q):
    """Randomised estimate of ||Q^T Q - I||_F for a row-major m x n Q, from
    REORTH_PROBES fixed-seed Gaussian columns (deterministic for a given Q)."""
    m, n = Q.shape
    key = (n, Q.device)
    if key not in _probe_cache:
        g = torch.Generator(device=Q.device)
        g.manual_seed(0x5EED)
        _probe_cache[key] = torch.randn(n, REORTH_PROBES, dtype=F64, device=Q.device, generator=g)
    X = _probe_cache[key]
    Y = _gemm_ksplit(Q, X)                              # Q X         (m x p)
    Z = _gemm_ksplit(Q.t(), Y, X.clone(), -1.0)          # Q^T Q X - X (n x p)
    return float(np.linalg.norm(Z.cpu().numpy())) / REORTH_PROBES ** 0.5


def _gemm_ksplit(A, B, C=None, beta=0.0, chunks=16):
    """C = A @ B + beta C for a skinny B (few columns) and a long K: the K
    range is cut into `chunks` slices computed as one grouped launch into a
    workspace, then summed by a one-row GEMM (a plain launch would give each
    of its few output tiles the whole K loop on one CU)."""
    m, k = A.shape
    p = B.shape[1]
    kc = (k + chunks - 1) // chunks
    parts = [(j * kc, min(k, (j + 1) * kc)) for j in range(chunks) if j * kc < k]
    W = torch.empty((len(parts), m, p), dtype=F64, device=A.device)
    gemm_many([(A[:, a:b], B[a:b, :], W[j], 1.0, 0.0) for j, (a, b) in enumerate(parts)], tile=0)
    if C is None:
        C = torch.empty((m, p), dtype=F64, device=A.device)
        beta = 0.0
    ones = torch.ones((1, len(parts)), dtype=F64, device=A.device)
    gemm(ones, W.view(len(parts), m * p), C.view(1, m * p), 1.0, beta)
    return C


def cholqr2(A, tolrel=None, mixed=False):
    """Reduced QR of a tall row-major matrix by two CholeskyQR passes (the
    second through cholqr_refine).  Returns (Q (m x n), R (n x n upper))
    with A = Q R.  mixed: the Q products on the FP32 matrix cores."""
    Q1, L1, _ = cholqr(A, tolrel, mixed)
    return _second_pass(Q1, L1, tolrel, mixed)


def _second_pass(Q1, L1, tolrel, mixed):
    Q, L2, _ = cholqr_refine(Q1, tolrel, thresh=CHOLQR_REFINE_THRESH_MIXED if mixed else CHOLQR_REFINE_THRESH,
                             mixed=mixed)
    R = gemm(L2.t(), L1.t(), tri=4 | 8)  # R = R2 R1 = L2^T L1^T, both upper
    return Q, R


def cholqr2_adaptive(A, allowance, tolrel=None):
    """CholeskyQR whose second pass runs only where it is needed: after the
    first pass, the loss of orthogonality ||Q1^T Q1 - I||_F is estimated
    (orthogonality_estimate); at or below `allowance` the first pass is kept.
    Returns (Q, R, spent) with A = Q R and spent the estimate of the kept
    first pass (0.0 when the second pass ran).  FP64 only."""
    Q1, L1, _ = cholqr(A, tolrel)
    est = orthogonality_estimate(Q1)
    if est <= allowance:
        return Q1, L1.t(), est
    Q, R = _second_pass(Q1, L1, tolrel, False)
    return Q, R, 0.0


def jacobi_tol(length):
    """Orthogonality threshold |x_p . x_q| <= tol ||x_p|| ||x_q||: the
    rounding error bound of a dot product of `length` terms, length * eps."""
    return max(JACOBI_TOL, float(length) * EPS)


def jacobi_blocked(X, gram=False, tol=None, max_sweeps=JACOBI_SWEEPS):
    """Block one-sided Jacobi on the rows of a row-major nv x len matrix (in
    place).  Returns (rt (nv x nv), s (nv,), sweeps) as jacobi_many does."""
    nv, ln = X.shape
    if tol is None:
        tol = jacobi_tol(ln)
    px, ldx, tr = as_matrix(X)
    if tr:
        raise ValueError("jacobi needs a row-major matrix")
    rt = torch.empty((nv, nv), dtype=F64, device=X.device)
    s = torch.empty(nv, dtype=F64, device=X.device)
    work = torch.empty(int(_lib.load().syn_jacobi_blocked_workspace(nv, ln)), dtype=F64, device=X.device)
    sweeps = ctypes.c_int(0)
    call("syn_jacobi_blocked_f64", px, nv, ln, ldx, _ptr(rt), nv, _ptr(s), _ptr(work), float(tol),
         int(max_sweeps), int(gram), ctypes.byref(sweeps), _stream())
    return rt, s, sweeps.value


def jacobi_gram_blocked(G, tol=None, max_sweeps=JACOBI_SWEEPS):
    """Two-sided block Jacobi eigen-decomposition of one symmetric Gram matrix
    (row-major nv x nv, diagonalised in place).  Returns (rt, s, sweeps) with
    rt rows the eigenvectors by decreasing eigenvalue and s = sqrt(lambda)."""
    nv, ln = G.shape
    if nv != ln:
        raise ValueError("jacobi_gram_blocked needs a square matrix")
    if tol is None:
        tol = jacobi_tol(nv)
    px, ldx, tr = as_matrix(G)
    if tr:
        raise ValueError("jacobi_gram_blocked needs a row-major matrix")
    rt = torch.empty((nv, nv), dtype=F64, device=G.device)
    s = torch.empty(nv, dtype=F64, device=G.device)
    it = np.zeros(1, dtype=_lib.GBJ_ITEM)
    it["g"], it["ldg"], it["rt"], it["ldr"], it["s"], it["nv"] = px, ldx, _ptr(rt), nv, _ptr(s), nv
    nva = np.array([nv], dtype=np.int64)
    work = torch.empty(int(_lib.load().syn_jacobi_gram_blocked_workspace(nva.ctypes.data_as(ctypes.c_void_p), 1)),
                       dtype=F64, device=G.device)
    counters = torch.zeros(1, dtype=torch.int32, device=G.device)
    dev_items = torch.empty(it.nbytes, dtype=torch.uint8, device=G.device)
    sweeps = ctypes.c_int(0)
    call("syn_jacobi_gram_blocked_f64", _ptr(dev_items), it.ctypes.data_as(ctypes.c_void_p), 1, _ptr(work),
         _ptr(counters), float(tol), int(max_sweeps), ctypes.byref(sweeps), _stream())
    return rt, s, sweeps.value
