"""ctypes binding of libsyngular_amd.so (the gfx950 C ABI in include/syngular_amd.h).

The shared object is built in-tree by ``__graft_entry__.build()`` (or ``make``
in syngular_amd/csrc).  There is no fallback: if the library is missing or no
GPU is visible, every device operation raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsyngular_amd.so")

SYN_OK = 0
SYN_EINVAL = -1
SYN_EHIP = -2
SYN_ENOTSUP = -3

APPLY_BLOCK = 1024  # SYN_APPLY_BLOCK

_P = np.uint64  # device pointer
_I = np.int64

# numpy mirrors of the item structs (align=True reproduces the C layout)
GEMM_ITEM = np.dtype(
    [("A", _P), ("B", _P), ("C", _P), ("m", _I), ("n", _I), ("k", _I), ("lda", _I), ("ldb", _I), ("ldc", _I),
     ("alpha", np.float64), ("beta", np.float64), ("transA", np.int32), ("transB", np.int32),
     ("block_start", _I), ("tiles_n", _I), ("tri", np.int32)], align=True)
APPLY_ITEM = np.dtype(
    [("x", _P), ("o", _P), ("p", _P), ("chil", _I), ("din", _I), ("chir", _I), ("dl", _I), ("dout", _I),
     ("dr", _I), ("block_start", _I)], align=True)
MPO_MPO_ITEM = np.dtype(
    [("a", _P), ("b", _P), ("p", _P), ("al", _I), ("din", _I), ("dmid", _I), ("ar", _I), ("bl", _I),
     ("dout", _I), ("br", _I), ("block_start", _I)], align=True)
DSUM_ITEM = np.dtype(
    [("a", _P), ("b", _P), ("out", _P), ("la", _I), ("ra", _I), ("lb", _I), ("rb", _I), ("mid", _I),
     ("lo", _I), ("ro", _I), ("edge", np.int32), ("block_start", _I)], align=True)
QR_ITEM = np.dtype(
    [("a", _P), ("rs", _I), ("cs", _I), ("m", _I), ("n", _I), ("keep", _I), ("q", _P), ("ldq", _I),
     ("q_trans", np.int32), ("work", _P), ("tau", _P)], align=True)
JACOBI_ITEM = np.dtype(
    [("x", _P), ("ldx", _I), ("nv", _I), ("len", _I), ("rt", _P), ("ldr", _I), ("s", _P), ("sweeps", _P),
     ("gram", np.int32)], align=True)
TRUNC_ITEM = np.dtype(
    [("s", _P), ("n", _I), ("chi_max", _I), ("cutoff", np.float64), ("keep", _P), ("discarded", _P)],
    align=True)
ROWSCALE_ITEM = np.dtype(
    [("x", _P), ("rows", _I), ("cols", _I), ("ld", _I), ("s", _P), ("mode", np.int32), ("block_start", _I)],
    align=True)

BJAC_ITEM = np.dtype(
    [("x", _P), ("ldx", _I), ("nv", _I), ("len", _I), ("rt", _P), ("ldr", _I), ("s", _P), ("block_start", _I),
     ("counter", _P), ("active", np.int32), ("gram", np.int32)], align=True)
EYE_ITEM = np.dtype([("x", _P), ("n", _I), ("ld", _I), ("block_start", _I)], align=True)
COPY_ITEM = np.dtype([("src", _P), ("dst", _P), ("n", _I), ("block_start", _I)], align=True)
DIAGMAX_ITEM = np.dtype([("g", _P), ("n", _I), ("ld", _I), ("out", _P), ("dropped", _P)], align=True)
CHOL_ITEM = np.dtype(
    [("g", _P), ("ldg", _I), ("l", _P), ("ldl", _I), ("x", _P), ("ldx", _I), ("n", _I), ("maxdiag", _P),
     ("tolrel", np.float64), ("dropped", _P)], align=True)

GBJ_ITEM = np.dtype(
    [("g", _P), ("ldg", _I), ("rt", _P), ("ldr", _I), ("s", _P), ("w", _P), ("nv", _I), ("solve_start", _I),
     ("tile_start", _I), ("counter", _P), ("active", np.int32)], align=True)

LUSIGN_ITEM = np.dtype(
    [("a", _P), ("lda", _I), ("bs", _I), ("free_col", _I), ("d", _P), ("linv", _P), ("ldl", _I), ("uinv", _P),
     ("ldu", _I), ("flag", _P)], align=True)
COLSCALE_ITEM = np.dtype(
    [("src", _P), ("lds", _I), ("dst", _P), ("ldd", _I), ("rows", _I), ("cols", _I), ("s", _P),
     ("trans", np.int32), ("block_start", _I)], align=True)

VECNORM_ITEM = np.dtype(
    [("x", _P), ("rows", _I), ("cols", _I), ("ld", _I), ("axis", np.int32), ("out", _P), ("block_start", _I)],
    align=True)

ITEM_DTYPES = [GEMM_ITEM, APPLY_ITEM, MPO_MPO_ITEM, DSUM_ITEM, QR_ITEM, JACOBI_ITEM, TRUNC_ITEM, ROWSCALE_ITEM,
               BJAC_ITEM, EYE_ITEM, COPY_ITEM, DIAGMAX_ITEM, CHOL_ITEM, GBJ_ITEM, LUSIGN_ITEM, COLSCALE_ITEM,
               VECNORM_ITEM]

# every symbol the header declares, with its ctypes signature
_VP, _I32, _I64, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
SIGNATURES = {
    "syn_last_error": (ctypes.c_char_p, []),
    "syn_version": (_I32, []),
    "syn_device_count": (_I32, [ctypes.POINTER(ctypes.c_int)]),
    "syn_struct_sizes": (_I32, [ctypes.POINTER(ctypes.c_int64), _I32]),
    "syn_gemm_f64": (_I32, [_VP, _I32, _I64, _I32, _VP]),
    "syn_gemm_mixed_f64": (_I32, [_VP, _I32, _I64, _I32, _VP]),
    "syn_gemm_bf16_f64": (_I32, [_VP, _I32, _I64, _I32, _VP]),
    "syn_mpo_apply_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_mpo_mpo_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_kron_bonds_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_direct_sum_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_qr_f64": (_I32, [_VP, _I32, _VP]),
    "syn_qr_sized_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_lu_sign_leaf_f64": (_I32, [_VP, _I32, _D, _VP]),
    "syn_hh_panel_f64": (_I32, [_VP, _I64, _VP, _I64, _VP, _VP, _I64, _I64, _I64, _I64, _VP]),
    "syn_colscale_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_jacobi_f64": (_I32, [_VP, _I32, _D, _I32, _VP]),
    "syn_jacobi_gram_f64": (_I32, [_VP, _I32, _I64, _D, _I32, _VP]),
    "syn_truncate_f64": (_I32, [_VP, _I32, _VP]),
    "syn_rowscale_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_vecnorms_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_jacobi_blocked_workspace": (_I64, [_I64, _I64]),
    "syn_jacobi_blocked_f64": (_I32, [_VP, _I64, _I64, _I64, _VP, _I64, _VP, _VP, _D, _I32, _I32,
                                      ctypes.POINTER(ctypes.c_int), _VP]),
    "syn_cholqr_workspace": (_I64, [_I64]),
    "syn_potrf_f64": (_I32, [_VP, _I64, _I64, _VP, _I64, _VP, _D, _VP, _VP]),
    "syn_cholqr_f64": (_I32, [_VP, _I64, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _D, _VP, _I32, _VP]),
    "syn_cholqr_noalloc_f64": (_I32, [_VP, _I64, _I64, _I64, _VP, _I64, _VP, _I64, _D, _VP, _I32, _VP]),
    "syn_cholqr_noalloc_max_n": (_I64, []),
    "syn_jacobi_blocked_many_f64": (_I32, [_VP, _VP, _I32, _VP, _VP, _D, _I32, ctypes.POINTER(ctypes.c_int), _VP]),
    "syn_jacobi_blocked_many_workspace": (_I64, [_VP, _VP, _I32]),
    "syn_eye_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_copy_f64": (_I32, [_VP, _I32, _I64, _VP]),
    "syn_diagmax_f64": (_I32, [_VP, _I32, _VP]),
    "syn_cholqr_refine_f64": (_I32, [_VP, _I64, _I64, _I64, _VP, _I64, _VP, _I64, _VP, _D, _D, _VP,
                                     ctypes.POINTER(ctypes.c_int), _I32, _VP]),
    "syn_jacobi_gram_blocked_f64": (_I32, [_VP, _VP, _I32, _VP, _VP, _D, _I32, ctypes.POINTER(ctypes.c_int), _VP]),
    "syn_jacobi_gram_blocked_workspace": (_I64, [_VP, _I32]),
    "syn_chol_leaf_f64": (_I32, [_VP, _I32, _VP]),
    "syn_prof_begin": (_I32, []),
    "syn_prof_select": (_I32, [ctypes.c_uint32]),
    "syn_prof_note_flops": (_I32, [_D]),
    "syn_prof_end": (_I32, [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), _I32]),
}

PROF_NTAGS = 20
PROF_TAGS = {
    0: "gemm_f64_single<64,64>",
    1: "gemm_f64_single<128,128>",
    2: "gemm_f64_grouped<64,64>",
    3: "gemm_f64_grouped<128,128>",
    4: "bjacobi_step_kernel",
    5: "chol_diag_kernel",
    6: "qr_kernel",
    7: "jacobi_kernel",
    8: "apply_kernel",
    9: "tjf_step_kernel",
    10: "tjf_step_bytes",
    11: "gemm_mixed<64,64>",
    12: "gemm_mixed<128,128>",
    13: "gemm_f64_splitk<128,128>",
    14: "gemm_f64_splitk<64,64>",
    15: "gemm_f64_pair<64,64>",
    16: "gemm_f64_pair<128,128>",
    17: "chol_node_kernel",
    18: "gemm_bf16<64,64>",
    19: "gemm_bf16<128,128>",
}


def prof_begin(tags=None):
    """Start event timing of the launches whose kernel name is in `tags`
    (every instrumented kernel when None)."""
    if tags is None:
        mask = 0xFFFFFFFF
    else:
        names = {v: k for k, v in PROF_TAGS.items()}
        unknown = [t for t in tags if t not in names]
        if unknown:
            raise ValueError(f"unknown kernel tags {unknown}; known: {sorted(names)}")
        mask = 0
        for t in tags:
            mask |= 1 << names[t]
    call("syn_prof_select", mask)
    call("syn_prof_begin")


def prof_end():
    """Stop profiling; returns {tag_name: (launches, ms, flops)} for the tags
    that recorded launches or work (a work-only tag has no launches)."""
    n = PROF_NTAGS
    la = (ctypes.c_int64 * n)()
    ms = (ctypes.c_double * n)()
    fl = (ctypes.c_double * n)()
    rc = load().syn_prof_end(la, ms, fl, n)
    if rc < 0:
        raise SyngularError(f"syn_prof_end failed ({rc}): {load().syn_last_error().decode()}")
    return {PROF_TAGS.get(i, f"tag{i}"): (la[i], ms[i], fl[i]) for i in range(n) if la[i] or fl[i]}


class SyngularError(RuntimeError):
    """A C-ABI call returned an error code."""


_lib = None


def load():
    """Load the shared object (no GPU needed) and bind every declared symbol."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SyngularError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()' or make -C syngular_amd/csrc)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check_abi():
    """Compare the C struct sizes with the numpy mirrors; raises on mismatch."""
    lib = load()
    n = lib.syn_struct_sizes(None, 0)
    buf = (ctypes.c_int64 * n)()
    lib.syn_struct_sizes(buf, n)
    sizes = list(buf)
    mine = [dt.itemsize for dt in ITEM_DTYPES]
    if sizes != mine:
        raise SyngularError(f"ABI mismatch: C sizes {sizes} vs Python {mine}")
    return sizes


def call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc is None:
        return rc
    if rc != SYN_OK:
        msg = load().syn_last_error().decode(errors="replace")
        if rc == SYN_EINVAL:
            raise ValueError(f"{name}: {msg}")
        raise SyngularError(f"{name} failed ({rc}): {msg}")
    return rc
