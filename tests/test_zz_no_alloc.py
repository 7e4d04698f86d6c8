"""The allocation-free CholeskyQR entry point (syn_cholqr_noalloc_f64)
against syn_cholqr_f64 with a caller workspace, through the C ABI: the same
kernels on the same inputs, so Q, L and the dropped-pivot count agree
bitwise.  Two calls in a row, the second smaller (the device scratch keeps
the first call's contents), then the size limit's error."""
import ctypes

import numpy as np
import pytest
import torch

from syngular_amd import _lib, ops

pytestmark = pytest.mark.gpu


def _run(fn, A, with_work):
    lib = _lib.load()
    m, n = A.shape
    Q = torch.full((m, n), np.nan, dtype=torch.float64, device=A.device)
    L = torch.full((n, n), np.nan, dtype=torch.float64, device=A.device)
    dropped = torch.full((1,), -7, dtype=torch.int32, device=A.device)
    args = [A.data_ptr(), m, n, n, Q.data_ptr(), n, L.data_ptr(), n]
    if with_work:
        work = torch.empty(int(lib.syn_cholqr_workspace(n)), dtype=torch.float64, device=A.device)
        args.append(work.data_ptr())
    args += [ops.cholqr_tol(m, n), dropped.data_ptr(), 0, ops._stream()]
    _lib.call(fn, *args)
    torch.cuda.synchronize()
    return Q.cpu().numpy(), L.cpu().numpy(), int(dropped.item())


@pytest.mark.parametrize("shapes", [((3000, 1000), (700, 300))])
def test_cholqr_noalloc_matches_workspace_entry(shapes):
    lib = _lib.load()
    nmax = int(lib.syn_cholqr_noalloc_max_n())
    assert nmax >= 1024
    rng = np.random.default_rng(11)
    for m, n in shapes:
        print(f"[gpu case] cholqr noalloc {m}x{n}", flush=True)
        A = torch.from_numpy(rng.standard_normal((m, n)) / np.sqrt(m)).to(ops.device())
        q0, l0, d0 = _run("syn_cholqr_f64", A, True)
        q1, l1, d1 = _run("syn_cholqr_noalloc_f64", A, False)
        assert d0 == d1 == 0
        assert np.array_equal(q0, q1) and np.array_equal(l0, l1)
        Ah = A.cpu().numpy()
        assert np.abs(q1 @ l1.T - Ah).max() <= 1e-13 * np.abs(Ah).max() * n
    A = torch.zeros((nmax + 8, nmax + 1), dtype=torch.float64, device=ops.device())
    with pytest.raises(ValueError, match="syn_cholqr_noalloc_f64"):
        _run("syn_cholqr_noalloc_f64", A, False)
