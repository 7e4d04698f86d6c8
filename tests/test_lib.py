"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every
symbol include/syngular_amd.h declares, and its struct layouts match the
ctypes mirrors.  No compute call is made."""
import os
import re

import pytest

from syngular_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "syngular_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(syn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_python_binds():
    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_loads_and_exports_every_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.syn_version() >= 100


def test_struct_layouts_match():
    sizes = _lib.check_abi()
    assert len(sizes) == len(_lib.ITEM_DTYPES)


def test_invalid_arguments_are_rejected_without_gpu():
    # argument validation happens on the host before any HIP call
    with pytest.raises(ValueError):
        _lib.call("syn_gemm_f64", None, -1, 0, 0, None)
    with pytest.raises(ValueError):
        _lib.call("syn_jacobi_f64", None, 1, -1.0, 10, None)
    assert "syn_jacobi_f64" in _lib.load().syn_last_error().decode()


def test_product_path_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "syngular_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\S+)", src, flags=re.M), f


def test_gram_workspace_matches_the_host_emulator():
    # the blocked Jacobi's workspace size is host arithmetic; the emulator the
    # host tests run on must ask for the same size at orders on both sides of
    # the wide-block switch and at its edges
    import ctypes

    import numpy as np

    from abi_emulator import _gram_workspace

    lib = _lib.load()
    for nv in ([1], [16], [17], [191, 192, 193], list(range(1, 700, 13)), [4096]):
        a = np.asarray(nv, np.int64)
        got = int(lib.syn_jacobi_gram_blocked_workspace(a.ctypes.data_as(ctypes.c_void_p), len(a)))
        assert got == _gram_workspace(a), nv
