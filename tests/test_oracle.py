"""Pins for the CPU oracle (oracle/).

Each check below is independent of the oracle's own formulas: the reference's
printed known-answer pairs (test/test_mps.py, test/test_mpo.py,
test/test_syn.py) turned into asserts, dense closed forms, Eckart-Young,
Schmidt values of the dense tensor and the Pythagoras identity of sequential
orthogonal projections.
"""
import numpy as np
import pytest

from oracle import reference as R
from oracle import rounding as RD


@pytest.fixture(autouse=True)
def _compress_default():
    R.MPO.COMPRESS = True
    yield
    R.MPO.COMPRESS = True


# --------------------------------------------------------------------------
# dense closed forms used as independent checks
# --------------------------------------------------------------------------
def proj(A):
    """Orthogonal projector onto range(A) (pseudo-inverse form, no QR)."""
    return A @ np.linalg.pinv(A)


def dense_decompose_left(T, bonds):
    """Left QR-truncation TT of a dense tensor: at step k project the
    (d_0..d_k, rest) unfolding onto the span of its first b_k columns."""
    shape = T.shape
    P = T.reshape(shape[0], -1)
    for k, b in enumerate(bonds):
        rows = int(np.prod(shape[: k + 1]))
        M = P.reshape(rows, -1)
        M = proj(M[:, :b]) @ M
        P = M
    return P.reshape(shape)


def dense_qr_truncate(sites, dim):
    """Reference >> (strict, left) from the ORIGINAL cores: the left block of
    bond k is projected onto the span of its first ``dim`` columns, then the
    untouched core k+1 is attached."""
    A = sites[0].reshape(-1, sites[0].shape[-1])
    for k in range(len(sites) - 1):
        A = proj(A[:, :dim]) @ A
        nxt = sites[k + 1]
        A = (A @ nxt.reshape(nxt.shape[0], -1)).reshape(-1, nxt.shape[-1])
    return A


def dense_sequential_truncation(T, kept):
    """TT-rounding in dense form: right to left, truncate the Schmidt
    decomposition at bond k to kept[k] terms (np.linalg.svd)."""
    shape = T.shape
    n = len(shape)
    for k in range(n - 1, 0, -1):
        M = T.reshape(int(np.prod(shape[:k])), -1)
        U, S, Vt = np.linalg.svd(M, full_matrices=False)
        c = kept[k]
        T = ((U[:, :c] * S[:c]) @ Vt[:c]).reshape(shape)
    return T


def random_sites(rng, dims, bonds):
    b = [1] + list(bonds) + [1]
    return [rng.standard_normal((b[k], d, b[k + 1])) for k, d in enumerate(dims)]


# --------------------------------------------------------------------------
# known-answer pairs from the reference's test scripts
# --------------------------------------------------------------------------
def test_mps_add_arange():
    # test/test_mps.py:51-70
    x = np.arange(8).reshape((2, 2, 2))
    X = R.MPS(x, bond_shape=(2, 2)).decompose()
    Y = R.MPS(x.copy(), bond_shape=(2, 2)).decompose()
    Z = X + Y
    np.testing.assert_allclose(Z[1, 1, 1], 14.0, rtol=1e-12)
    np.testing.assert_allclose(Z.to_tensor(), 2 * x, atol=1e-12)
    np.testing.assert_allclose(R.retrieve_all(Z), 2 * x, atol=1e-12)


def test_mps_dot_and_normalize():
    # test/test_mps.py:100-126
    x = np.arange(4 ** 3).reshape((4, 4, 4)).astype(float)
    X = R.MPS(x, bond_shape=(4, 4)).decompose()
    np.testing.assert_allclose(X | X, np.sum(x * x), rtol=1e-12)
    np.testing.assert_allclose(X.dot(), np.linalg.norm(x), rtol=1e-12)
    X.normalize()
    np.testing.assert_allclose(X.dot(), 1.0, rtol=1e-12)


@pytest.mark.parametrize("mode", ["left", "right"])
def test_mps_decompose_arange_exact(mode):
    # test/test_mps.py:129-157 (rank-2 unfoldings, bonds 2 -> exact)
    x = np.arange(4 ** 3).reshape((4, 4, 4)).astype(float)
    X = R.MPS(x, bond_shape=(2, 2)).decompose(mode=mode)
    np.testing.assert_allclose(X.to_tensor(), x, atol=1e-9)


def test_mpo_decompose_elements():
    # test/test_mpo.py:44-64
    w = np.arange(2 ** 4).reshape((2, 2, 2, 2))
    W = R.MPO(w, bond_shape=(4,)).decompose()
    np.testing.assert_allclose(W[(1, 1), (1, 1)], w[1, 1, 1, 1], rtol=1e-12)
    z = np.arange(4 ** 3 * 3 ** 3).reshape((4, 4, 4, 3, 3, 3))
    Z = R.MPO(z, bond_shape=(3, 3)).decompose()
    np.testing.assert_allclose(Z[(3, 1, 0), (1, 2, 1)], z[3, 1, 0, 1, 2, 1], rtol=1e-10)
    np.testing.assert_allclose(Z.to_tensor(), z, atol=1e-8)


def test_mpo_orthogonality():
    # test/test_mpo.py:66-112
    w = np.arange(4 ** 6).reshape((4,) * 6)
    W = R.MPO(w, bond_shape=(2, 2)).decompose()
    W.left_orthonormalization()
    np.testing.assert_allclose(np.diag(W.left_orthogonality(0)), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.diag(W.left_orthogonality(1)), 1.0, atol=1e-12)
    W.right_orthonormalization()
    np.testing.assert_allclose(np.diag(W.right_orthogonality(1)), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.diag(W.right_orthogonality(2)), 1.0, atol=1e-12)
    np.testing.assert_allclose(W[(1, 0, 0), (1, 0, 0)], w[1, 0, 0, 1, 0, 0], rtol=1e-9)


def test_mpo_dot():
    # test/test_mpo.py:160-175
    x = np.arange(2 ** 2).reshape((2, 2))
    X = R.MPS(x, bond_shape=(2,)).decompose()
    np.testing.assert_allclose(X | X, 14.0, rtol=1e-12)


def test_mpo_matmul_uncompressed():
    # test/test_mpo.py:177-196 with the implicit truncation disabled
    w = np.arange(2 ** 4).reshape((2, 2, 2, 2))
    x = np.arange(2 ** 2).reshape((2, 2))
    W = R.MPO(w, bond_shape=(2,)).decompose()
    X = R.MPS(x, bond_shape=(2,)).decompose()
    R.MPO.COMPRESS = False
    Z = (W @ W) @ X
    np.testing.assert_allclose(Z.to_tensor().reshape(4), [1680, 1940, 2200, 2460], rtol=1e-12)


def test_mpo_add_chain():
    # test/test_mpo.py:256-282
    x = np.arange(1, 17).reshape((2, 2, 2, 2))
    y = np.arange(18, 34).reshape((2, 2, 2, 2))
    X = R.MPO(x, bond_shape=(3,)).decompose()
    Y = R.MPO(y, bond_shape=(3,)).decompose()
    Z = X + Y + Y + Y + Y + Y + Y + Y + Y + Y
    z = x + 9 * y
    np.testing.assert_allclose(Z[(1, 1), (1, 1)], z[1, 1, 1, 1], rtol=1e-10)
    np.testing.assert_allclose(Z.to_tensor(), z, atol=1e-8)


def test_mpo_diag_right_orthonormalization():
    # test/test_mpo.py:296-327
    y = np.zeros((4, 4))
    np.fill_diagonal(y, [1, 2, 3, 4])
    y = y.reshape((2, 2, 2, 2))
    Y = R.MPO(y, bond_shape=(4,)).decompose().right_orthonormalization()
    assert np.max(np.abs(y - Y.to_tensor())) < 1e-12


def test_syn_mul():
    # test/test_syn.py:10-27
    x = np.arange(2 ** 2)
    X = R.MPS(x.reshape((2, 2)), bond_shape=(2,)).decompose()
    w = np.arange(2 ** 4).reshape((4, 4))
    W = R.MPO(w.reshape((2, 2, 2, 2)), bond_shape=(2,)).decompose()
    np.testing.assert_allclose(R.mul(W, X).to_tensor().reshape(4), x @ w, rtol=1e-12)
    np.testing.assert_allclose(R.mul(X, W).to_tensor().reshape(4), x @ w, rtol=1e-12)
    np.testing.assert_allclose(R.mul(X, X), x @ x, rtol=1e-12)


# --------------------------------------------------------------------------
# dense closed forms on random inputs
# --------------------------------------------------------------------------
def test_decompose_projector_closed_form():
    rng = np.random.default_rng(1)
    T = rng.standard_normal((3, 4, 5, 3))
    bonds = (2, 5, 2)
    X = R.MPS(T, bond_shape=bonds).decompose()
    np.testing.assert_allclose(X.to_tensor(), dense_decompose_left(T, bonds), atol=1e-12)
    # an MPO decomposes its interleaved (i0,o0,i1,o1,..) tensor the same way
    t = rng.standard_normal((2, 3, 2, 2, 2, 3))
    W = R.MPO(t, bond_shape=(3, 2)).decompose()
    ti = np.transpose(t, (0, 3, 1, 4, 2, 5)).reshape(4, 6, 6)
    ref = dense_decompose_left(ti, (3, 2)).reshape(2, 2, 3, 2, 2, 3).transpose(0, 2, 4, 1, 3, 5)
    np.testing.assert_allclose(W.to_tensor(), ref, atol=1e-12)


def test_qr_truncation_projector_closed_form():
    rng = np.random.default_rng(2)
    sites = random_sites(rng, (3, 4, 3, 2), (6, 7, 5))
    X = R.MPS.from_sites(sites)
    Z = X >> 3
    np.testing.assert_allclose(Z.to_tensor().reshape(-1), dense_qr_truncate(sites, 3).reshape(-1), atol=1e-12)
    # stale bond_shape of the returned copy, updated shape list
    assert Z.bond_shape == (6, 7, 5)
    assert [s[2] for s in Z.shape[:-1]] == [3, 3, 3]
    # dim >= min(bond_shape) returns self unchanged
    assert (X >> 5) is X


def test_mpo_rshift_mutates_receiver():
    rng = np.random.default_rng(3)
    sites = [rng.standard_normal(s) for s in [(1, 2, 2, 5), (5, 2, 2, 4), (4, 2, 2, 1)]]
    W = R.MPO.from_sites(sites)
    Z = W >> 2
    assert W.bond_shape == (2, 2)          # overwritten on the receiver
    assert Z.bond_shape == (5, 4)          # stale on the result
    assert Z.sites is W.sites              # aliasing copy
    assert Z.sites[0].shape == (1, 2, 2, 2)


def test_rshift_too_small_unfolding_raises():
    # (lrank*d) < dim cannot be reshaped back, the reference raises ValueError
    rng = np.random.default_rng(4)
    sites = random_sites(rng, (2, 2, 2), (16, 16))
    with pytest.raises(ValueError):
        R.MPS.from_sites(sites) >> 4


def test_apply_dense_identity():
    rng = np.random.default_rng(5)
    xs = random_sites(rng, (3, 2, 4), (3, 2))
    os_ = [rng.standard_normal(s) for s in [(1, 3, 2, 2), (2, 2, 3, 3), (3, 4, 2, 1)]]
    X = R.MPS.from_sites(xs)
    O = R.MPO.from_sites(os_)
    R.MPO.COMPRESS = False
    Y = O @ X
    x = X.to_tensor()
    o = O.to_tensor()  # (i0,i1,i2,o0,o1,o2)
    y = np.einsum("abc,abcdef->def", x, o)
    np.testing.assert_allclose(Y.to_tensor(), y, atol=1e-12)
    # MPO @ MPO is the matrix product in the (in, out) view
    os2 = [rng.standard_normal(s) for s in [(1, 3, 3, 2), (2, 2, 2, 2), (2, 4, 4, 1)]]
    O2 = R.MPO.from_sites(os2)
    os3 =[rng.standard_normal(s) for s in [(1, 3, 3, 2), (2, 2, 2, 2), (2, 4, 4, 1)]]
    O3 = R.MPO.from_sites(os3)
    P = O2 @ O3
    a = O2.to_tensor().reshape(24, 24)
    b = O3.to_tensor().reshape(24, 24)
    np.testing.assert_allclose(P.to_tensor().reshape(24, 24), a @ b, atol=1e-12)
    assert R.mul(O3, O2).to_tensor().shape == (3, 2, 4, 3, 2, 4)
    np.testing.assert_allclose(R.mul(O3, O2).to_tensor().reshape(24, 24), a @ b, atol=1e-12)


def test_mpo_hadamard_dense():
    # tensor/matrix_product_operator.py:126-161 with the implicit truncation off
    rng = np.random.default_rng(12)
    a = R.MPO.from_sites([rng.standard_normal(s) for s in [(1, 2, 3, 2), (2, 2, 2, 1)]])
    b = R.MPO.from_sites([rng.standard_normal(s) for s in [(1, 2, 3, 3), (3, 2, 2, 1)]])
    R.MPO.COMPRESS = False
    np.testing.assert_allclose((a * b).to_tensor(), a.to_tensor() * b.to_tensor(), atol=1e-12)


def test_add_dense():
    rng = np.random.default_rng(6)
    A = R.MPS.from_sites(random_sites(rng, (2, 3, 2), (2, 3)))
    B = R.MPS.from_sites(random_sites(rng, (2, 3, 2), (4, 2)))
    np.testing.assert_allclose((A + B).to_tensor(), A.to_tensor() + B.to_tensor(), atol=1e-12)
    assert (A + B).bond_shape == (6, 5)
    R.MPO.COMPRESS = False
    oa = R.MPO.from_sites([rng.standard_normal(s) for s in [(1, 2, 3, 2), (2, 2, 2, 1)]])
    ob = R.MPO.from_sites([rng.standard_normal(s) for s in [(1, 2, 3, 3), (3, 2, 2, 1)]])
    np.testing.assert_allclose((oa + ob).to_tensor(), oa.to_tensor() + ob.to_tensor(), atol=1e-12)


def test_overlap_dense_and_elementwise():
    rng = np.random.default_rng(7)
    A = R.MPS.from_sites(random_sites(rng, (3, 2, 4, 2), (3, 5, 2)))
    B = R.MPS.from_sites(random_sites(rng, (3, 2, 4, 2), (2, 4, 3)))
    a, b = A.to_tensor(), B.to_tensor()
    np.testing.assert_allclose(R.retrieve_all(A), a, atol=1e-12)
    np.testing.assert_allclose(A | B, np.sum(a * b), rtol=1e-12)
    np.testing.assert_allclose(A.left_orthogonality(1), A._left(A.sites[1], 1).T @ A._left(A.sites[1], 1))


def test_left_orthonormalization_keeps_tensor():
    rng = np.random.default_rng(8)
    A = R.MPS.from_sites(random_sites(rng, (3, 4, 4, 3), (3, 5, 3)))
    t = A.to_tensor()
    A.left_orthonormalization()
    np.testing.assert_allclose(A.to_tensor(), t, atol=1e-12)
    for k in range(3):
        np.testing.assert_allclose(A.left_orthogonality(k), np.eye(A.shape[k][2]), atol=1e-12)
    A.right_orthonormalization()
    np.testing.assert_allclose(A.to_tensor(), t, atol=1e-12)
    for k in range(1, 4):
        np.testing.assert_allclose(A.right_orthogonality(k), np.eye(A.shape[k][0]), atol=1e-12)


# --------------------------------------------------------------------------
# TT-rounding (SVD form)
# --------------------------------------------------------------------------
def test_keep_count_rule():
    s = np.array([4.0, 2.0, 1.0, 0.5])
    w = s ** 2
    assert RD.keep_count(s) == 4
    assert RD.keep_count(s, chi_max=2) == 2
    # tail after 3 kept = 0.25 ; relative = 0.25/21.25
    assert RD.keep_count(s, cutoff=0.25 / w.sum()) == 3
    assert RD.keep_count(s, cutoff=0.25 / w.sum() * 0.999) == 4
    assert RD.keep_count(s, cutoff=1.0) == 1
    assert RD.keep_count(np.zeros(3), cutoff=1e-3) == 1


def test_round_eckart_young_two_sites():
    rng = np.random.default_rng(9)
    sites = random_sites(rng, (6, 5), (4,))
    cores, sv, kept, disc = RD.tt_round(sites, chi_max=2)
    M = R.dense_mps(sites)
    U, S, Vt = np.linalg.svd(M)
    best = (U[:, :2] * S[:2]) @ Vt[:2]
    np.testing.assert_allclose(R.dense_mps(cores), best, atol=1e-12)
    np.testing.assert_allclose(sv[1], S[:4], rtol=1e-12)
    np.testing.assert_allclose(disc[1], np.sum(S[2:] ** 2), rtol=1e-12)


def test_round_dense_sequential_and_pythagoras():
    rng = np.random.default_rng(10)
    sites = random_sites(rng, (2, 3, 2, 3, 2), (2, 6, 6, 3))
    T = R.dense_mps(sites)
    cores, sv, kept, disc = RD.tt_round(sites, chi_max=3, cutoff=1e-3)
    np.testing.assert_allclose(R.dense_mps(cores), dense_sequential_truncation(T, kept), atol=1e-11)
    # singular values at the first truncated bond are the Schmidt values of T
    S = np.linalg.svd(T.reshape(-1, 2), compute_uv=False)
    np.testing.assert_allclose(sv[4], S, rtol=1e-12)
    n0 = np.sum(T * T)
    n1 = RD.norm2(cores)
    np.testing.assert_allclose(n0 - n1, sum(disc[1:]), rtol=1e-9, atol=1e-12 * n0)
    # the result is right-orthonormal from core 1 on
    for k in range(1, 5):
        c = cores[k].reshape(cores[k].shape[0], -1)
        np.testing.assert_allclose(c @ c.T, np.eye(c.shape[0]), atol=1e-12)


def test_apply_and_round_exact_when_untruncated():
    rng = np.random.default_rng(11)
    xs = random_sites(rng, (2, 2, 2, 2), (2, 3, 2))
    os_ = [rng.standard_normal(s) for s in [(1, 2, 2, 2), (2, 2, 2, 2), (2, 2, 2, 2), (2, 2, 2, 1)]]
    cores, sv, kept, disc = RD.apply_and_round(os_, xs)
    x = R.dense_mps(xs)
    o = R.dense_mpo(os_)
    y = np.einsum("abcd,abcdefgh->efgh", x, o)
    np.testing.assert_allclose(R.dense_mps(cores), y, atol=1e-11)
    assert all(d < 1e-20 for d in disc[1:])


# --------------------------------------------------------------------------
# non-strict compress (left and right modes) and MPO element access
# --------------------------------------------------------------------------
def dense_qr_truncate_right(sites, dim):
    """Right-mode QR truncation from the ORIGINAL cores, dense: at bond k
    (right to left) the right block is projected onto the span of its first
    ``dim`` rows, then core k-1 is attached."""
    B = sites[-1].reshape(sites[-1].shape[0], -1)
    for k in range(len(sites) - 1, 0, -1):
        B = (proj(B.T[:, :dim]) @ B.T).T
        prv = sites[k - 1]
        B = (prv.reshape(-1, prv.shape[-1]) @ B).reshape(prv.shape[0], -1)
    return B


def test_mps_compress_non_strict_left_and_right():
    """Non-strict compress = orthonormalize (tensor unchanged, pinned above)
    then the QR truncation whose dense effect is the projector closed form."""
    rng = np.random.default_rng(21)
    sites = random_sites(rng, (3, 4, 4, 3), (3, 6, 3))
    for mode, dense in (("left", dense_qr_truncate), ("right", dense_qr_truncate_right)):
        A = R.MPS.from_sites([s.copy() for s in sites])
        if mode == "left":
            A.left_orthonormalization()
        else:
            A.right_orthonormalization()
        expect = dense([s.copy() for s in A.sites], 2).reshape(-1)
        A.compress(2, mode=mode, strict=False)
        np.testing.assert_allclose(A.to_tensor().reshape(-1), expect, atol=1e-12)
        assert all(2 in (s.shape[0], s.shape[-1]) for s in A.sites)


def test_mpo_compress_non_strict_and_retrieve():
    rng = np.random.default_rng(22)
    w = [1, 3, 5, 1]
    cores = [rng.standard_normal((w[k], 2, 3, w[k + 1])) for k in range(3)]
    M = R.MPO.from_sites([c.copy() for c in cores])
    t = M.to_tensor()
    # element access = product of the sliced core matrices
    for idx in [((0, 1, 1), (2, 0, 1)), ((1, 0, 0), (0, 2, 2))]:
        v = np.ones((1, 1))
        for k in range(3):
            v = v @ cores[k][:, idx[0][k], idx[1][k], :]
        assert abs(M[idx] - v[0, 0]) < 1e-12
        assert abs(t[idx[0] + idx[1]] - v[0, 0]) < 1e-12
    M.left_orthonormalization()
    lo = [c.reshape(c.shape[0], -1, c.shape[-1]).copy() for c in M.sites]
    expect = dense_qr_truncate(lo, 2).reshape(-1)
    M.compress(2, strict=False)
    got = R.dense_mps([c.reshape(c.shape[0], -1, c.shape[-1]) for c in M.sites]).reshape(-1)
    np.testing.assert_allclose(got, expect, atol=1e-12)


# --------------------------------------------------------------------------
# MPS.apply (tensor/matrix_product_state.py:487-534) and the qubit register
# (quantum/qbit.py, quantum/circuit.py)
# --------------------------------------------------------------------------
from oracle import quantum as QO  # noqa: E402


def dense_gate(T, G, qubits):
    """Apply the 2m-leg gate G (outputs first) to axes `qubits` of T."""
    m = len(qubits)
    out = np.tensordot(G, T, axes=(list(range(m, 2 * m)), list(qubits)))
    return np.moveaxis(out, list(range(m)), list(qubits))


def statevector(size, structure):
    """Dense simulation of a gate list as the reference's register applies
    it (quantum/qbit.py:30-61): (G, i) acts on qubits i..i+m-1; (G, a, b)
    on neighbours acts on (min, max) whatever the order; on qubits further
    apart it acts on (a, b), a the first leg, and when a < b the SWAPs are
    undone from the top, which leaves qubits a and b exchanged."""
    psi = np.zeros((2,) * size)
    psi[(0,) * size] = 1.0
    for op in structure:
        G = op[0]
        m = G.ndim // 2
        if len(op) == 2:
            psi = dense_gate(psi, G, list(range(op[1], op[1] + m)))
            continue
        a, b = op[1], op[2]
        if abs(a - b) == 1:
            psi = dense_gate(psi, G, [min(a, b), max(a, b)])
        else:
            psi = dense_gate(psi, G, [a, b])
            if a < b:
                psi = np.swapaxes(psi, a, b)
    return psi


def test_mps_apply_cx_known_answer():
    # test/test_mps.py:9-19 (printed there): CX on sites 1, 2 of arange(8)
    x = np.arange(8.0).reshape(2, 2, 2)
    X = R.MPS(x, bond_shape=(2, 2)).decompose()
    before = [s.copy() for s in X.sites]
    Y = X.apply(QO.CX, 1)
    want = x.copy()
    want[:, 1, :] = x[:, 1, ::-1]          # flip qubit 2 where qubit 1 is set
    np.testing.assert_allclose(Y.to_tensor(), want, atol=1e-12)
    assert Y.shape == X.shape
    for a, b in zip(X.sites, before):      # the receiver is untouched
        np.testing.assert_array_equal(a, b)


def test_mps_apply_one_site_exact():
    rng = np.random.default_rng(21)
    sites = random_sites(rng, (2, 3, 2, 4), (2, 5, 3))
    X = R.MPS.from_sites(sites)
    G = rng.standard_normal((3, 3))
    Y = X.apply(G, 1)
    np.testing.assert_allclose(Y.to_tensor(), dense_gate(X.to_tensor(), G, [1]), atol=1e-12)
    assert Y.shape == X.shape


def test_mps_apply_two_site_projector_closed_form():
    # the split keeps the span of the first `rank` columns of the block's
    # left unfolding: block' = P B with P = B_r B_r^+ (no QR involved)
    rng = np.random.default_rng(22)
    dims, bonds = (2, 3, 2, 3, 2), (2, 4, 3, 2)
    sites = random_sites(rng, dims, bonds)
    X = R.MPS.from_sites(sites)
    G = rng.standard_normal((3, 2, 3, 2))          # sites 1, 2 have d = 3, 2
    Y = X.apply(G, 1)
    blk = np.einsum("asb,btc->astc", sites[1], sites[2])       # (2, 3, 2, 3)
    blk = np.einsum("stuv,auvc->astc", G, blk)
    B = blk.reshape(2 * 3, 2 * 3)
    B = proj(B[:, :bonds[1]]) @ B                  # rank 4 of 6: a real truncation
    new = [sites[0], B.reshape(2, 3, 2, 3)] + sites[3:]
    T = new[0]
    for s in new[1:]:
        T = np.tensordot(T, s, axes=([T.ndim - 1], [0]))
    np.testing.assert_allclose(Y.to_tensor(), T.reshape(dims), atol=1e-12)
    assert Y.shape == X.shape
    # bond large enough for the block: the gate is applied exactly
    sites = random_sites(rng, (2, 2, 2), (2, 2))
    Z = R.MPS.from_sites(sites)
    U = rng.standard_normal((2, 2, 2, 2))
    np.testing.assert_allclose(Z.apply(U, 0).to_tensor(), dense_gate(Z.to_tensor(), U, [0, 1]), atol=1e-12)


def test_mps_apply_three_site_reshape_rule():
    # three sites at bond 2 with r_1 = d_2 r_2: the reference's reshape
    # (rank, -1, right bond of the next site) is consistent, the gate exact
    rng = np.random.default_rng(23)
    sites = random_sites(rng, (2, 2, 2), (2, 2))
    X = R.MPS.from_sites(sites)
    G = rng.standard_normal((2,) * 6)
    np.testing.assert_allclose(X.apply(G, 0).to_tensor(), dense_gate(X.to_tensor(), G, [0, 1, 2]), atol=1e-12)
    # a middle bond that is not d_2 r_2 cannot be reshaped back: ValueError
    sites = random_sites(rng, (2, 2, 2, 2), (2, 3, 2))
    with pytest.raises(ValueError):
        R.MPS.from_sites(sites).apply(G, 0)


def test_mps_apply_errors():
    rng = np.random.default_rng(24)
    X = R.MPS.from_sites(random_sites(rng, (2, 2, 2), (2, 2)))
    with pytest.raises(UnboundLocalError):
        X.apply(QO.CX, 0, strict=False)
    with pytest.raises(ValueError):
        X.apply(np.ones((2, 2, 2)), 0)


def test_qbit_ground_and_bell():
    q = QO.Qbit(3)
    np.testing.assert_allclose(q.to_tensor(), np.eye(8)[0], atol=0)
    # test/test_quantum.py:155-164: H on 0 then CX(0, 1)
    b = QO.Qbit(2) @ (QO.H, 0) @ (QO.CX, 0, 1)
    np.testing.assert_allclose(b.to_tensor(), np.array([1, 0, 0, 1]) / np.sqrt(2), atol=1e-12)


def test_qbit_binary_and_toffoli():
    q = QO.Qbit.from_binary("110")
    QO.Qbit.LSB = False
    try:
        assert q.to_binary() == "110"
        assert (q @ (QO.TOFFOLI, 0)).to_binary() == "111"
        assert (QO.Qbit.from_binary("010") @ (QO.TOFFOLI, 0)).to_binary() == "010"
    finally:
        QO.Qbit.LSB = True
    assert QO.Qbit.from_binary("110").to_binary() == "011"


def test_qbit_ghz_and_bernstein_vazirani_dense():
    # GHZ: Schmidt rank 2 on every cut, so bond 2 holds it exactly
    ghz = [(QO.H, 0)] + [(QO.CX, i, i + 1) for i in range(4)]
    np.testing.assert_allclose(QO.run_circuit(5, ghz).to_tensor(), statevector(5, ghz).reshape(-1), atol=1e-12)
    # test/test_quantum.py:82-113 with a short token: CX(ancilla, i)
    token = "1011"
    n = len(token)
    bv = [(QO.H, n), (QO.Z, n)] + [(QO.H, i) for i in range(n)]
    bv += [(QO.CX, n, i) for i in range(n) if token[i] == "1"] + [(QO.H, i) for i in range(n)]
    got = QO.run_circuit(n + 1, bv).to_tensor()
    want = statevector(n + 1, bv).reshape(-1)
    np.testing.assert_allclose(got, want, atol=1e-12)
    # only the neighbour of the ancilla sees the phase kick-back (its CX has
    # the query qubit as control): the register ends as |0001> (x) |->
    np.testing.assert_allclose(got[[2, 3]], [2 ** -0.5, -2 ** -0.5], atol=1e-12)


def test_qbit_swap_network_quirk():
    # (G, a, b) with a < b not neighbours: the SWAPs are undone from the top
    # (quantum/qbit.py:88-103), which exchanges qubits a and b afterwards
    prep = [(QO.H, 0), (QO.X, 1)]
    q = QO.run_circuit(3, prep) @ (QO.CX, 0, 2)
    want = dense_gate(statevector(3, prep), QO.CX, [0, 2]).transpose(2, 1, 0)
    np.testing.assert_allclose(q.to_tensor(), want.reshape(-1), atol=1e-12)
    # with the control above the target the network is undone correctly
    q = QO.run_circuit(3, [(QO.X, 2), (QO.H, 0)]) @ (QO.CX, 2, 0)
    want = dense_gate(statevector(3, [(QO.X, 2), (QO.H, 0)]), QO.CX, [2, 0])
    np.testing.assert_allclose(q.to_tensor(), want.reshape(-1), atol=1e-12)


# --------------------------------------------------------------------------
# DMRG environment blocks (variational/dmrg.py:65-112) and restore (:219-239)
# --------------------------------------------------------------------------
from oracle import variational as VO  # noqa: E402


def random_mpo_sites(rng, dims, bonds):
    b = [1] + list(bonds) + [1]
    return [rng.standard_normal((b[k], d, d, b[k + 1])) for k, d in enumerate(dims)]


def dense_sandwich(x, w_sites):
    """sum_{s,t} x[s] W[in=t, out=s] x[t] from the dense operator."""
    W = R.dense_mpo(w_sites)                      # (in..., out...)
    n = x.ndim
    Wm = W.reshape(int(np.prod(x.shape)), -1)     # rows: inputs t, cols: outputs s
    v = x.reshape(-1)
    return float(v @ Wm.T @ v) if n else 0.0


@pytest.mark.parametrize("dims,bonds,wb", [((2, 2, 2, 2, 2), (2, 4, 4, 2), (3, 2, 3, 2)),
                                           ((3, 2, 4, 2, 3, 2), (3, 5, 6, 4, 2), (2, 4, 3, 3, 2))])
def test_environment_blocks_pair_to_dense_form(dims, bonds, wb):
    rng = np.random.default_rng(sum(dims) + sum(bonds))
    xs = random_sites(rng, dims, bonds)
    ws = random_mpo_sites(rng, dims, wb)
    E = dense_sandwich(R.dense_mps(xs), ws)
    n = len(dims)
    Lb = VO.left_blocks(ws, xs)
    Rb = VO.right_blocks(ws, xs)
    assert [b is None for b in Lb] == [False] * (n - 2) + [True] * 2
    assert [b is None for b in Rb] == [True] * 2 + [False] * (n - 2)
    for k in range(2, n - 1):
        # left block of sites <= k-1 against right block of sites >= k
        assert abs(np.sum(Lb[k - 1] * Rb[k]) - E) <= 1e-10 * max(1.0, abs(E))
    # the outermost blocks closed with the remaining two sites by hand
    Z1 = VO._sandwich(xs[n - 2], ws[n - 2])
    Z2 = VO._sandwich(xs[n - 1], ws[n - 1])
    tail = np.einsum("ace,acebdf,bdfghi->", Lb[n - 3], Z1, Z2)
    assert abs(tail - E) <= 1e-10 * max(1.0, abs(E))
    assert Lb[0].shape == (bonds[0], wb[0], bonds[0]) and Rb[n - 1].shape == (bonds[-1], wb[-1], bonds[-1])


def test_restore_projector_closed_form():
    rng = np.random.default_rng(31)
    shapes = [(1, 2, 2), (2, 3, 4), (4, 2, 3), (3, 2, 1)]
    site = rng.standard_normal((2, 3, 2, 3))
    s1, s2 = VO.restore(shapes, site, 1)
    assert s1.shape == shapes[1] and s2.shape == shapes[2]
    B = site.reshape(6, 6)
    got = s1.reshape(6, 4) @ s2.reshape(4, 6)
    np.testing.assert_allclose(got, proj(B[:, :4]) @ B, atol=1e-12)
    with pytest.raises(Exception):
        VO.restore(shapes, site, 4)


def test_environment_block_legs_against_dense_halves():
    """A left block is the dense left half of ket, operator and bra with
    their physical legs contracted: the ket's with the operator's output leg
    and the bra's with its input leg (labels [1,3,2], [4,6,3,5], [7,6,8] at
    variational/dmrg.py:96-100), indices [ket, operator, bra]."""
    rng = np.random.default_rng(33)
    dims, bonds, wb = (2, 3, 2, 2), (2, 3, 2), (2, 3, 2)
    xs = random_sites(rng, dims, bonds)
    ws = random_mpo_sites(rng, dims, wb)
    # dense halves of sites 0..1
    A = np.tensordot(xs[0], xs[1], axes=([2], [0]))[0]            # (s0, s1, b)
    Wl = np.tensordot(ws[0], ws[1], axes=([3], [0]))[0]           # (t0, s0, t1, s1, d)
    want = np.einsum("xyb,uxvyd,uvf->bdf", A, Wl, A)              # ket s meets out, bra t meets in
    Lb = VO.left_blocks(ws, xs)
    np.testing.assert_allclose(Lb[1], want, atol=1e-12)
    # swapping the operator's in/out legs moves the result off (W is not symmetric)
    swapped = np.einsum("xyb,xuyvd,uvf->bdf", A, Wl, A)
    assert np.abs(swapped - Lb[1]).max() > 1e-3


# --------------------------------------------------------------------------
# MPO.apply on a sub-chain (tensor/matrix_product_operator.py:582-646)
# --------------------------------------------------------------------------
def random_mpo_rect(rng, din, dout, bonds):
    b = [1] + list(bonds) + [1]
    return [rng.standard_normal((b[k], din[k], dout[k], b[k + 1])) for k in range(len(din))]


def test_mpo_apply_single_site_dense():
    rng = np.random.default_rng(51)
    sites = random_mpo_rect(rng, (2, 3, 2), (3, 2, 2), (3, 4))
    W = R.MPO.from_sites(list(sites))
    B = rng.standard_normal((1, 2, 4, 1))              # site 1: output 2 -> 4
    before = R.dense_mpo(sites)                        # (i0, i1, i2, o0, o1, o2)
    assert W.apply(R.MPO.from_sites([B]), [1]) is None
    want = np.einsum("abcdef,ep->abcdpf", before, B[0, :, :, 0])
    np.testing.assert_allclose(R.dense_mpo(W.sites), want, atol=1e-12)
    assert W.sites[1].shape == (3, 3, 4, 4) and W.shape[1] == (3, 3, 2, 4)   # shapes not updated


def test_mpo_apply_two_sites_reshape_rule_exact():
    # r_j0 = l d_in d_out: the split is exact, so site j0 x site j1 is T in
    # its own memory order (l, in0, in1, out0, out1, r)
    rng = np.random.default_rng(52)
    sites = random_mpo_rect(rng, (2, 2, 2), (2, 2, 2), (4, 3))
    sites[1] = rng.standard_normal((4, 2, 2, 16))
    sites[2] = rng.standard_normal((16, 2, 2, 1))
    ops_ = random_mpo_rect(rng, (2, 2), (2, 2), (3,))
    T = np.einsum("aijb,cjpe,bklf,elqg->aikpqfcg", sites[1], ops_[0], sites[2], ops_[1])[..., 0, 0]
    W = R.MPO.from_sites(list(sites))
    W.apply(R.MPO.from_sites(list(ops_)), [1, 2])
    got = np.tensordot(W.sites[1], W.sites[2], axes=([3], [0]))
    np.testing.assert_allclose(got.reshape(-1), T.reshape(-1), atol=1e-10)


def test_mpo_apply_two_sites_projector_closed_form():
    rng = np.random.default_rng(53)
    sites = random_mpo_rect(rng, (2, 2, 2, 2), (2, 2, 2, 2), (3, 5, 3))
    ops_ = random_mpo_rect(rng, (2, 2), (2, 2), (2,))
    T = np.einsum("aijb,cjpe,bklf,elqg->aikpqfcg", sites[1], ops_[0], sites[2], ops_[1])[..., 0, 0]
    L = T.reshape(3 * 2 * 2, -1)                       # rows (l, in0, in1) in T's order
    want = proj(L[:, :5]) @ L                          # rank 5 of 12
    W = R.MPO.from_sites(list(sites))
    W.apply(R.MPO.from_sites(list(ops_)), [1, 2])
    got = W.sites[1].reshape(12, 5) @ W.sites[2].reshape(5, -1)
    np.testing.assert_allclose(got, want, atol=1e-10)


def test_mpo_apply_errors():
    rng = np.random.default_rng(54)
    sites = random_mpo_rect(rng, (2, 2, 2), (2, 2, 2), (3, 3))
    W = R.MPO.from_sites(list(sites))
    with pytest.raises(ValueError):
        W.apply(R.MPO.from_sites(random_mpo_rect(rng, (2, 2), (2, 2), (2,))), [0, 2])
    wide = [rng.standard_normal((2, 2, 2, 1))]         # operator left bond 2 cannot be squeezed
    with pytest.raises(ValueError):
        W.apply(R.MPO.from_sites(wide), [1])
    assert W.apply(np.eye(2), [0]) is None
    for a, b in zip(W.sites, sites):
        np.testing.assert_array_equal(a, b)
    D = R.MPO(rng.standard_normal((2, 2, 2, 2)), bond_shape=(2,))
    with pytest.raises(Exception):
        D.apply(R.MPO.from_sites(wide), [0])
