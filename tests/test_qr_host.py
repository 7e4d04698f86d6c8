"""Host logic of the fast Householder QR (ops.householder_q) through the C-ABI
emulator on the CPU: CholeskyQR2, the blocked sign LU over 64-column leaves
and the column scaling must reproduce numpy.linalg.qr (LAPACK signs), and the
fallbacks (rank-deficient block, ambiguous sign, keep > n) must route to the
one-work-group Householder kernel."""
import numpy as np
import pytest
import torch

import abi_emulator


@pytest.fixture
def emu(monkeypatch):
    from syngular_amd import ops

    # the host logic is checked with the path switched on
    monkeypatch.setattr(ops, "HH_FAST_PATH", True)
    return abi_emulator.install(monkeypatch, tensor_api=True)


def lapack_q(A, keep):
    return np.linalg.qr(A, mode="complete")[0][:, :keep]


@pytest.mark.parametrize("m,n,keep", [(300, 200, 150), (200, 200, 200), (256, 130, 130), (129, 64, 64),
                                      (700, 300, 300)])
def test_householder_q_matches_lapack(emu, m, n, keep):
    from syngular_amd import ops

    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n))
    Q = ops.householder_q(torch.from_numpy(A), keep).numpy()
    np.testing.assert_allclose(Q, lapack_q(A, keep), atol=1e-11)
    assert "syn_lu_sign_leaf_f64" in emu.calls and "syn_qr_f64" not in emu.calls
    # the transposed input and output of the right sweeps
    emu.calls.clear()
    At = torch.from_numpy(A.T.copy()).t()
    Qt = ops.householder_q(At, keep, q_trans=True).numpy()
    np.testing.assert_allclose(Qt, lapack_q(A, keep).T, atol=1e-11)
    assert "syn_qr_f64" not in emu.calls


def test_householder_q_fallbacks(emu):
    from syngular_amd import ops

    rng = np.random.default_rng(3)
    # rank-deficient leading block: the first pass drops a pivot
    A = rng.standard_normal((200, 100))
    A[:, 5] = A[:, 3]
    ops.householder_q(torch.from_numpy(A), 80)
    assert "syn_qr_f64" in emu.calls
    # structured input whose columns LAPACK leaves unreflected: ambiguous signs
    emu.calls.clear()
    ops.householder_q(torch.from_numpy(np.eye(128)[:, :96].copy()), 96)
    assert "syn_lu_sign_leaf_f64" in emu.calls and "syn_qr_f64" in emu.calls
    # fewer than 64 kept columns, or more kept than the matrix has
    for shape, keep in [((100, 40), 40), ((100, 70), 90)]:
        emu.calls.clear()
        ops.householder_q(torch.from_numpy(rng.standard_normal(shape)), keep)
        assert emu.calls == ["syn_qr_f64"]


@pytest.mark.parametrize("m,n,keep", [(300, 200, 150), (200, 200, 200), (129, 100, 100), (520, 700, 300),
                                      (100, 300, 120), (256, 70, 256)])
def test_blocked_qr_matches_lapack(emu, m, n, keep):
    """The blocked Householder QR (panels of 32, compact WY on the GEMM)
    reproduces numpy.linalg.qr's complete Q columns, both orientations,
    including more kept columns than the matrix has and wide matrices."""
    from syngular_amd import ops

    rng = np.random.default_rng(m + 3 * n + 7 * keep)
    A = rng.standard_normal((m, n))
    want = lapack_q(A, keep)
    Q = ops.qr_blocked(torch.from_numpy(A), keep).numpy()
    assert Q.shape == want.shape
    np.testing.assert_allclose(Q, want, atol=1e-11)
    Qt = ops.qr_blocked(torch.from_numpy(A.T.copy()).t(), keep, q_trans=True).numpy()
    np.testing.assert_allclose(Qt, want.T, atol=1e-11)
    Qs = ops.qr_blocked(torch.from_numpy(np.ascontiguousarray(A[:, : n - 3])), keep).numpy() if n > keep else None
    if Qs is not None:
        np.testing.assert_allclose(Qs, lapack_q(A[:, : n - 3], keep), atol=1e-11)
    assert "syn_hh_panel_f64" in emu.calls
