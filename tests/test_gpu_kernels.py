"""Kernel-level parity of the gfx950 library against NumPy / the oracle.

Floating-point tolerances (stated per test): FP64 kernels must agree with the
FP64 oracle to relative error <= 1e-12 on products and <= 1e-10 on
factorisation results (which depend on the summation order of long dot
products).
"""
import numpy as np
import pytest
import torch

from oracle import reference as oref
from oracle import rounding as ornd

pytestmark = pytest.mark.gpu


def dev():
    from syngular_amd import ops

    return ops.device()


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (130, 67, 129), (257, 300, 31), (512, 384, 700)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("tile", [0, 1])
def test_gemm_f64(m, n, k, ta, tb, tile):
    from syngular_amd import ops

    case(f"gemm m={m} n={n} k={k} ta={ta} tb={tb} tile={tile}")
    rng = np.random.default_rng(m * 131 + n * 7 + k + ta * 3 + tb)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    a, b = T(A), T(B)
    c = T(C0)
    ops.gemm_many([(a.t() if ta else a, b.t() if tb else b, c, 0.75, -0.5)], tile=tile)
    ref = 0.75 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.5 * C0
    got = c.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))) * np.sqrt(k)


def test_gemm_grouped_uneven():
    from syngular_amd import ops

    case("gemm grouped uneven")
    rng = np.random.default_rng(3)
    probs, refs = [], []
    for (m, n, k) in [(3, 9, 4), (100, 1, 17), (65, 65, 65), (1, 200, 50), (33, 129, 2)]:
        A, B = rng.standard_normal((m, k)), rng.standard_normal((k, n))
        C = T(np.zeros((m, n)))
        probs.append((T(A), T(B), C, 1.0, 0.0))
        refs.append(A @ B)
    ops.gemm_many(probs)
    for (_, _, C, _, _), r in zip(probs, refs):
        np.testing.assert_allclose(C.cpu().numpy(), r, rtol=1e-12, atol=1e-12)


def test_apply_and_mpo_mpo_and_kron():
    from syngular_amd import ops

    case("apply / mpo_mpo / kron")
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3, 4, 5))
    O = rng.standard_normal((2, 4, 3, 6))
    (P,) = ops.apply_many([(T(O), T(X))])
    np.testing.assert_allclose(P.cpu().numpy(), oref.apply_site(O, X), rtol=1e-13, atol=1e-13)
    A = rng.standard_normal((2, 3, 4, 3))
    B = rng.standard_normal((3, 4, 2, 2))
    (Q,) = ops.mpo_mpo_many([(T(A), T(B))])
    ref = np.einsum("aijb,cjpe->acipbe", A, B).reshape(6, 3, 2, 6)
    np.testing.assert_allclose(Q.cpu().numpy(), ref, rtol=1e-13, atol=1e-13)
    B2 = rng.standard_normal((2, 3, 4, 5))
    K = ops.kron_bonds(T(A), T(B2)).cpu().numpy()
    for i in range(3):
        for o in range(4):
            np.testing.assert_allclose(K[:, i, o, :], np.kron(A[:, i, o, :], B2[:, i, o, :]), rtol=1e-14)


def test_direct_sum_edges():
    from syngular_amd import ops

    case("direct sum")
    rng = np.random.default_rng(5)
    a0, b0 = rng.standard_normal((1, 3, 2)), rng.standard_normal((1, 3, 4))
    a1, b1 = rng.standard_normal((2, 3, 5)), rng.standard_normal((4, 3, 1))
    a2, b2 = rng.standard_normal((5, 3, 1)), rng.standard_normal((1, 3, 1))
    outs = ops.direct_sum_many([(T(a0), T(b0), 1), (T(a1), T(b1), 0), (T(a2), T(b2), 2)])
    A = oref.MPS.from_sites([a0, a1, a2])
    B = oref.MPS.from_sites([b0, b1, b2])
    for got, ref in zip(outs, (A + B).sites):
        np.testing.assert_array_equal(got.cpu().numpy(), ref)


@pytest.mark.parametrize("m,n,keep", [(6, 4, 4), (20, 7, 3), (9, 9, 9), (4, 16, 2), (300, 40, 40), (40, 3, 10),
                                      (1100, 64, 64)])
def test_householder_qr_matches_lapack(m, n, keep):
    from syngular_amd import ops

    case(f"qr m={m} n={n} keep={keep}")
    rng = np.random.default_rng(m + n + keep)
    L = rng.standard_normal((m, n))
    (Q,) = ops.qr_many([(T(L), keep)])
    Qn, Rn = np.linalg.qr(L, mode="complete")
    kq = min(keep, m)
    got = Q.cpu().numpy()
    assert got.shape == (m, kq)
    # same reflector convention as LAPACK: the columns agree, not only their span
    np.testing.assert_allclose(got, Qn[:, :kq], atol=1e-12)
    (Qt,) = ops.qr_many([(T(L), keep)], q_trans=True)
    np.testing.assert_allclose(Qt.cpu().numpy(), Qn[:, :kq].T, atol=1e-12)


def test_jacobi_gram_blocked_wide_batch():
    """A call with hundreds of Gram matrices: those of order 384 and above
    hold more than 2^23 entries together and take the 32-row blocks, the
    rest the fused 16-row steps after them.  Singular values against numpy,
    orthonormal eigenvector rows and the reconstruction of G, over orders
    with partial and odd blocks on both sides of the switch."""
    from syngular_amd import batch as B

    rng = np.random.default_rng(21)
    orders = [int(v) for v in rng.integers(17, 520, size=300)]
    orders[:8] = [17, 32, 33, 64, 383, 384, 385, 519]
    assert sum(g * g for g in orders if g >= 384) >= 1 << 23
    case(f"jacobi gram blocked wide, {len(orders)} problems, orders {min(orders)}..{max(orders)}")
    Cs = [rng.standard_normal((g + 3, g)) for g in orders]
    Cs[5] = Cs[5] * np.logspace(0, -3, orders[5])
    Gs = [c.T @ c for c in Cs]
    dev = torch.device("cuda")
    G = [torch.tensor(x, device=dev) for x in Gs]
    R = [torch.empty_like(x) for x in G]
    S = [torch.empty(x.shape[0], dtype=torch.float64, device=dev) for x in G]
    ptr = lambda ts: np.array([t.data_ptr() for t in ts], dtype=np.uint64)
    sweeps = B.jacobi_gram_blocked_table(ptr(G), np.array(orders, np.int64), ptr(R), ptr(S))
    torch.cuda.synchronize()
    assert 0 < sweeps < 60
    for q in range(0, len(orders), 7):
        g = orders[q]
        s, r = S[q].cpu().numpy(), R[q].cpu().numpy()
        ref = np.linalg.svd(Cs[q], compute_uv=False)
        assert np.all(np.diff(s) <= 0), q
        np.testing.assert_allclose(s, ref, rtol=1e-9, atol=1e-12 * ref[0], err_msg=str(q))
        np.testing.assert_allclose(r @ r.T, np.eye(g), atol=1e-12, err_msg=str(q))
        np.testing.assert_allclose(r.T @ np.diag(s ** 2) @ r, Gs[q], atol=1e-11 * ref[0] ** 2, err_msg=str(q))


def test_householder_qr_sized_launch_classes():
    """The grouped QR launches each LDS class (16, 48, 96 KB, and the
    workspace class beyond) with the LDS its largest matrix needs: a batch
    with matrices in every class gives LAPACK's Q for each."""
    from syngular_amd import batch as B

    shapes = [(40, 8), (64, 32), (120, 40), (200, 60), (150, 80), (400, 48), (1100, 64), (30, 29)]
    case(f"qr sized classes {shapes}")
    need = [m * n for m, n in shapes]
    classes = set(int(np.searchsorted(B.QR_LDS_CLASSES, v)) for v in need)
    assert classes == {0, 1, 2, 3}
    rng = np.random.default_rng(5)
    dev = torch.device("cuda")
    As = [rng.standard_normal(sh) for sh in shapes]
    A = [torch.tensor(x, device=dev) for x in As]
    Q = [torch.empty(sh, dtype=torch.float64, device=dev) for sh in shapes]
    work = [torch.empty(m * n, dtype=torch.float64, device=dev) for m, n in shapes]
    tau = [torch.empty(n, dtype=torch.float64, device=dev) for _, n in shapes]
    ptr = lambda ts: np.array([t.data_ptr() for t in ts], dtype=np.uint64)
    m = np.array([sh[0] for sh in shapes], np.int64)
    n = np.array([sh[1] for sh in shapes], np.int64)
    B.qr_table(ptr(A), m, n, ptr(Q), ptr(work), ptr(tau))
    torch.cuda.synchronize()
    for q, x in enumerate(As):
        Qn = np.linalg.qr(x, mode="reduced")[0]
        np.testing.assert_allclose(Q[q].cpu().numpy(), Qn, atol=1e-12, err_msg=str(shapes[q]))


def test_householder_qr_strided_and_rank_deficient():
    from syngular_amd import ops

    case("qr strided / rank deficient")
    rng = np.random.default_rng(6)
    L = rng.standard_normal((7, 12))
    (Q,) = ops.qr_many([(T(L).t(), 7)])  # QR of the transposed view
    Qn, _ = np.linalg.qr(L.T, mode="complete")
    np.testing.assert_allclose(Q.cpu().numpy(), Qn[:, :7], atol=1e-12)
    x = np.arange(64).reshape(4, 16).astype(float)  # rank 2
    (Q,) = ops.qr_many([(T(x), 4)])
    q = Q.cpu().numpy()
    np.testing.assert_allclose(q.T @ q, np.eye(4), atol=1e-12)
    np.testing.assert_allclose(q @ (q.T @ x), x, atol=1e-10)


@pytest.mark.parametrize("m,n,keep", [(512, 512, 256), (300, 200, 150), (200, 200, 200), (129, 100, 100),
                                      (520, 700, 300), (100, 300, 120), (256, 70, 256), (2048, 1024, 1024),
                                      (4096, 512, 512)])
def test_blocked_householder_qr_matches_lapack(m, n, keep):
    """Panels of 32 on syn_hh_panel_f64, compact-WY updates and Q on the
    GEMM: the same columns as numpy.linalg.qr (LAPACK signs), both
    orientations, including kept columns beyond n and wide matrices."""
    from syngular_amd import ops

    case(f"blocked qr m={m} n={n} keep={keep}")
    rng = np.random.default_rng(m * 5 + n * 11 + keep)
    A = rng.standard_normal((m, n))
    Qn = np.linalg.qr(A, mode="complete")[0][:, :min(keep, m)]
    Q = ops.qr_blocked(T(A), keep).cpu().numpy()
    assert Q.shape == Qn.shape
    np.testing.assert_allclose(Q, Qn, atol=1e-11)
    At = T(A.T.copy()).t()
    Qt = ops.qr_blocked(At, keep, q_trans=True).cpu().numpy()
    np.testing.assert_allclose(Qt, Qn.T, atol=1e-11)


def test_blocked_householder_qr_rank_deficient():
    from syngular_amd import ops

    case("blocked qr rank deficient / structured")
    x = np.arange(300 * 90, dtype=float).reshape(300, 90)        # rank 2
    Q = ops.qr_blocked(T(x), 80).cpu().numpy()
    np.testing.assert_allclose(Q.T @ Q, np.eye(80), atol=1e-12)
    np.testing.assert_allclose(Q @ (Q.T @ x), x, atol=1e-8 * np.abs(x).max())
    E = np.eye(200)[:, :120].copy()                              # columns LAPACK leaves unreflected
    Q = ops.qr_blocked(T(E), 120).cpu().numpy()
    np.testing.assert_allclose(Q, np.linalg.qr(E, mode="complete")[0][:, :120], atol=1e-12)


def test_cholqr_split_k_gram():
    """An 8192 x 4096 CholeskyQR pass: its Gram matrix (528 live tiles of 128)
    takes the split-K launch; Q is orthonormal and Q L^T = A."""
    from syngular_amd import ops

    case("cholqr 8192 x 4096 (split-K Gram)")
    g = torch.Generator(device="cpu").manual_seed(5)
    A = (torch.randn(8192, 4096, generator=g, dtype=torch.float64) / 64.0).to(dev())
    Q, L, dropped = ops.cholqr(A)
    assert int(dropped.item()) == 0
    G = ops.gemm(Q.t(), Q)
    I = torch.eye(4096, dtype=torch.float64, device=dev())
    assert float((G - I).abs().max()) < 1e-10           # one pass: cond(A)^2 eps
    R = ops.gemm(Q, L.t())
    assert float((R - A).abs().max()) <= 1e-12 * float(A.abs().max()) * 64
    Q2, R2 = ops.cholqr2(A)
    G2 = ops.gemm(Q2.t(), Q2)
    assert float((G2 - I).abs().max()) < 1e-13


FAST_QR_SKIP = ("the CholeskyQR2 + sign-LU QR path is switched off (ops.HH_FAST_PATH): its first device run "
                "faulted the GPU and the cause is not identified")


@pytest.mark.skip(reason=FAST_QR_SKIP)
@pytest.mark.parametrize("m,n,keep", [(512, 512, 256), (300, 200, 150), (200, 200, 200), (129, 64, 64),
                                      (2048, 1024, 1024), (1000, 700, 640)])
def test_householder_q_fast_path_matches_lapack(m, n, keep):
    """CholeskyQR2 + blocked sign LU + column scaling: the same columns as
    numpy.linalg.qr (LAPACK signs), in both output orientations."""
    from syngular_amd import ops

    case(f"householder_q m={m} n={n} keep={keep}")
    rng = np.random.default_rng(m * 7 + n * 3 + keep)
    A = rng.standard_normal((m, n))
    Qn = np.linalg.qr(A, mode="complete")[0][:, :keep]
    Q = ops.householder_q(T(A), keep).cpu().numpy()
    assert Q.shape == Qn.shape
    np.testing.assert_allclose(Q, Qn, atol=1e-11)
    At = T(A.T.copy()).t()                           # transposed view, as in the right sweeps
    Qt = ops.householder_q(At, keep, q_trans=True).cpu().numpy()
    np.testing.assert_allclose(Qt, Qn.T, atol=1e-11)


@pytest.mark.skip(reason=FAST_QR_SKIP)
def test_householder_q_fallbacks_on_device():
    from syngular_amd import ops

    case("householder_q fallbacks")
    rng = np.random.default_rng(12)
    A = rng.standard_normal((200, 100))
    A[:, 5] = A[:, 3]                                # rank-deficient block: first pass drops a pivot
    Q = ops.householder_q(T(A), 80).cpu().numpy()
    np.testing.assert_allclose(Q, np.linalg.qr(A, mode="complete")[0][:, :80], atol=1e-11)
    E = np.eye(128)[:, :96].copy()                   # unreflected columns: ambiguous signs
    Q = ops.householder_q(T(E), 96).cpu().numpy()
    np.testing.assert_allclose(Q, np.linalg.qr(E, mode="complete")[0][:, :96], atol=1e-12)


@pytest.mark.parametrize("nv,ln", [(2, 3), (5, 8), (16, 40), (33, 33), (64, 200), (100, 120)])
def test_jacobi_svd(nv, ln):
    from syngular_amd import ops

    case(f"jacobi nv={nv} len={ln}")
    rng = np.random.default_rng(nv * 100 + ln)
    X0 = rng.standard_normal((nv, ln))
    x = T(X0)
    rt, s, sw = ops.jacobi_many([x])[0]
    s_ref = np.linalg.svd(X0, compute_uv=False)
    k = min(nv, ln)
    np.testing.assert_allclose(s.cpu().numpy()[:k], s_ref, rtol=1e-12)
    R = rt.cpu().numpy()
    np.testing.assert_allclose(R @ R.T, np.eye(nv), atol=1e-12)
    np.testing.assert_allclose(R.T @ x.cpu().numpy(), X0, atol=1e-11)
    assert 0 < int(sw.item()) < 60


def test_jacobi_gram_mode():
    from syngular_amd import ops

    case("jacobi gram")
    rng = np.random.default_rng(8)
    C = rng.standard_normal((50, 12))
    G = T(C.T @ C)
    rt, s, _ = ops.jacobi_many([G], gram=True)[0]
    np.testing.assert_allclose(s.cpu().numpy(), np.linalg.svd(C, compute_uv=False), rtol=1e-11)
    V = rt.cpu().numpy().T
    U = C @ V / s.cpu().numpy()
    np.testing.assert_allclose(U.T @ U, np.eye(12), atol=1e-10)


def test_truncate_matches_oracle_rule():
    from syngular_amd import ops

    case("truncate")
    rng = np.random.default_rng(9)
    cases = []
    for n in (1, 5, 64, 65, 200):
        s = np.sort(np.exp(rng.uniform(-12, 0, n)))[::-1].copy()
        for chi, cut in [(None, 0.0), (3, 0.0), (None, 1e-6), (10, 1e-3), (None, 0.5), (1000, 1e-12)]:
            cases.append((s, chi, cut))
    keep, disc = ops.truncate_many([(T(s), chi, cut) for s, chi, cut in cases])
    keep, disc = keep.cpu().numpy(), disc.cpu().numpy()
    for i, (s, chi, cut) in enumerate(cases):
        c = ornd.keep_count(s, chi, cut)
        assert keep[i] == c, (len(s), chi, cut, keep[i], c)
        np.testing.assert_allclose(disc[i], np.sum(s[c:] ** 2), rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("m,n", [(40, 40), (200, 64), (700, 300), (1100, 513)])
def test_cholqr2(m, n):
    from syngular_amd import ops

    case(f"cholqr2 m={m} n={n}")
    rng = np.random.default_rng(m + n)
    A0 = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -3, n))
    Q, R = ops.cholqr2(T(A0))
    q, r = Q.cpu().numpy(), R.cpu().numpy()
    np.testing.assert_allclose(q.T @ q, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(q @ r, A0, atol=1e-13 * np.linalg.norm(A0, 2))
    assert np.allclose(np.tril(r, -1), 0.0)
    # R agrees with LAPACK's up to the signs of its rows
    _, rn = np.linalg.qr(A0)
    np.testing.assert_allclose(np.abs(r), np.abs(rn), atol=1e-11 * np.abs(rn).max())


def test_cholqr_rank_deficient_drops():
    from syngular_amd import ops

    case("cholqr rank deficient")
    rng = np.random.default_rng(21)
    B = rng.standard_normal((300, 40))
    A0 = np.concatenate([B, B[:, :10] @ rng.standard_normal((10, 24))], axis=1)  # rank 40 of 64
    Q, L, dropped = ops.cholqr(T(A0))
    assert int(dropped.item()) == 24
    Q2, L2, d2 = ops.cholqr(Q.clone())
    q = Q2.cpu().numpy()
    r = (L.cpu().numpy() @ L2.cpu().numpy()).T
    np.testing.assert_allclose(q @ r, A0, atol=1e-10 * np.abs(A0).max())
    kept = np.linalg.norm(q, axis=0) > 0.5
    assert kept.sum() == 40
    np.testing.assert_allclose(q[:, kept].T @ q[:, kept], np.eye(40), atol=1e-12)


@pytest.mark.parametrize("n,dep", [(65, ()), (100, ()), (128, ()), (100, range(70, 100)),
                                   (128, tuple(range(40, 52)) + tuple(range(110, 128)))])
def test_cholqr_node_sizes(n, dep):
    """Unfoldings of 65..128 columns, where the recursive Cholesky-inverse is
    one node (two 64-column leaves and the products between them); dep:
    columns made combinations of earlier ones, which the factorisation drops
    (in the second leaf, or in both)."""
    from syngular_amd import ops

    case(f"cholqr node n={n} dropped={len(dep)}")
    rng = np.random.default_rng(n + len(dep))
    A0 = rng.standard_normal((400, n)) @ np.diag(np.logspace(0, -2, n))
    for j in dep:
        A0[:, j] = A0[:, :j] @ rng.standard_normal(j) / np.sqrt(j)
    Q, L, dropped = ops.cholqr(T(A0))
    assert int(dropped.item()) == len(dep)
    Q2, L2, d2 = ops.cholqr(Q.clone())
    q = Q2.cpu().numpy()
    r = (L.cpu().numpy() @ L2.cpu().numpy()).T
    np.testing.assert_allclose(q @ r, A0, atol=1e-10 * np.abs(A0).max())
    kept = np.linalg.norm(q, axis=0) > 0.5
    assert kept.sum() == n - len(dep)
    np.testing.assert_allclose(q[:, kept].T @ q[:, kept], np.eye(n - len(dep)), atol=1e-12)
    if not dep:
        # full rank: the factor matches LAPACK's R up to the signs of its rows
        _, rn = np.linalg.qr(A0)
        np.testing.assert_allclose(np.abs(r), np.abs(rn), atol=1e-10 * np.abs(rn).max())


@pytest.mark.parametrize("nv,ln,gram", [(17, 40, False), (100, 130, False), (300, 300, True), (512, 512, True)])
def test_jacobi_blocked(nv, ln, gram):
    from syngular_amd import ops

    case(f"jacobi blocked nv={nv} len={ln} gram={gram}")
    rng = np.random.default_rng(nv + ln)
    if gram:
        C = rng.standard_normal((2 * nv, nv)) @ np.diag(np.logspace(0, -2, nv))
        X0 = C.T @ C
        s_ref = np.linalg.svd(C, compute_uv=False)
    else:
        X0 = rng.standard_normal((nv, ln))
        s_ref = np.linalg.svd(X0, compute_uv=False)
    rt, s, sweeps = ops.jacobi_blocked(T(X0), gram=gram)
    k = min(nv, ln)
    np.testing.assert_allclose(s.cpu().numpy()[:k], s_ref, rtol=1e-11 if not gram else 1e-9)
    R = rt.cpu().numpy()
    np.testing.assert_allclose(R @ R.T, np.eye(nv), atol=1e-12)
    assert 0 < sweeps < 60


@pytest.mark.parametrize("tri", [1, 2, 4, 8, 16, 4 | 8, 2 | 16])
@pytest.mark.parametrize("n", [64, 200, 300])
def test_gemm_structural_zero_hints(tri, n):
    from syngular_amd import ops

    case(f"gemm tri={tri} n={n}")
    rng = np.random.default_rng(tri * 1000 + n)
    A = rng.standard_normal((n, n))
    B = rng.standard_normal((n, n))
    if tri & 2:
        A = np.tril(A)
    if tri & 4:
        A = np.triu(A)
    if tri & 8:
        B = np.triu(B)
    if tri & 16:
        B = np.tril(B)
    C = T(np.zeros((n, n)))
    ops.gemm_many([(T(A), T(B), C, 1.0, 0.0, tri)], tile=1 if n > 128 else 0)
    got, ref = C.cpu().numpy(), A @ B
    if tri & 1:
        got, ref = np.tril(got), np.tril(ref)
    np.testing.assert_allclose(got, ref, atol=1e-12 * n)


@pytest.mark.parametrize("g,rank", [(1, 1), (2, 2), (7, 7), (33, 20), (64, 64), (64, 10)])
def test_jacobi_gram_two_sided(g, rank):
    """Two-sided LDS Jacobi of C^T C: eigenvalues = squared singular values of
    C, rows of rt orthonormal eigenvectors sorted decreasingly, input kept."""
    from syngular_amd import ops

    case(f"jacobi gram two-sided g={g} rank={rank}")
    rng = np.random.default_rng(g * 100 + rank)
    C = rng.standard_normal((g + 5, rank)) @ rng.standard_normal((rank, g))
    G0 = C.T @ C
    G = T(G0)
    (rt, s), = ops.jacobi_gram_many([G])
    s = s.cpu().numpy()
    R = rt.cpu().numpy()
    ref = np.linalg.svd(C, compute_uv=False)
    assert np.all(np.diff(s) <= 0)
    np.testing.assert_allclose(s[:rank], ref[:rank], rtol=1e-10, atol=1e-12 * ref[0])
    assert np.all(s[rank:] <= 1e-6 * ref[0])
    np.testing.assert_allclose(R @ R.T, np.eye(g), atol=1e-12)
    np.testing.assert_allclose(R.T @ np.diag(s ** 2) @ R, G0, atol=1e-11 * ref[0] ** 2)
    np.testing.assert_array_equal(G.cpu().numpy(), G0)


def test_jacobi_gram_rejects_oversize():
    from syngular_amd import ops

    with pytest.raises(ValueError):
        ops.jacobi_gram_many([T(np.eye(65))])


@pytest.mark.parametrize("g,rank", [(65, 65), (100, 100), (257, 257), (300, 120), (512, 512)])
def test_jacobi_gram_blocked(g, rank):
    """Two-sided block Jacobi of a Gram matrix: eigenvalues = squared
    singular values, orthonormal sorted eigenvector rows, reconstruction."""
    from syngular_amd import ops

    case(f"jacobi gram blocked g={g} rank={rank}")
    rng = np.random.default_rng(g + rank)
    C = rng.standard_normal((g + 7, rank)) @ rng.standard_normal((rank, g))
    G0 = C.T @ C
    rt, s, sweeps = ops.jacobi_gram_blocked(T(G0))
    s = s.cpu().numpy()
    R = rt.cpu().numpy()
    ref = np.linalg.svd(C, compute_uv=False)
    assert np.all(np.diff(s) <= 0) and 0 < sweeps < 60
    np.testing.assert_allclose(s[:rank], ref[:rank], rtol=1e-9, atol=1e-12 * ref[0])
    assert np.all(s[rank:] <= 1e-6 * ref[0])
    np.testing.assert_allclose(R @ R.T, np.eye(g), atol=1e-12)
    np.testing.assert_allclose(R.T @ np.diag(s ** 2) @ R, G0, atol=1e-11 * ref[0] ** 2)


def test_jacobi_gram_blocked_error_budget_tolerance():
    """The blocked Jacobi stopped at the error budget's threshold (every
    |G_pq| <= 1e-8 sqrt(G_pp G_qq)): no more sweeps than at the FP64 one,
    singular values within 2e-8 (a near-degenerate pair may move by tol
    lambda), eigenvector rows orthonormal to rounding (a product of
    rotations) and G reconstructed within the residual bound g tol s_max^2."""
    from syngular_amd import batch as B

    rng = np.random.default_rng(23)
    orders = [40, 96, 129, 256, 300]
    case(f"jacobi gram blocked at tol {B.TOL_BUDGET:g}, orders {orders}")
    Cs = [rng.standard_normal((g + 3, g)) for g in orders]
    Gs = [c.T @ c for c in Cs]
    dev = torch.device("cuda")
    ptr = lambda ts: np.array([t.data_ptr() for t in ts], dtype=np.uint64)
    sweeps = {}
    for tol in (None, B.TOL_BUDGET):
        G = [torch.tensor(x, device=dev) for x in Gs]
        R = [torch.empty_like(x) for x in G]
        S = [torch.empty(x.shape[0], dtype=torch.float64, device=dev) for x in G]
        sweeps[tol] = B.jacobi_gram_blocked_table(ptr(G), np.array(orders, np.int64), ptr(R), ptr(S), tol)
        torch.cuda.synchronize()
    assert 0 < sweeps[B.TOL_BUDGET] <= sweeps[None]
    for q, g in enumerate(orders):
        s, r = S[q].cpu().numpy(), R[q].cpu().numpy()
        ref = np.linalg.svd(Cs[q], compute_uv=False)
        assert np.all(np.diff(s) <= 0), q
        np.testing.assert_allclose(s, ref, rtol=2e-8, atol=1e-12 * ref[0], err_msg=str(q))
        np.testing.assert_allclose(r @ r.T, np.eye(g), atol=1e-12, err_msg=str(q))
        np.testing.assert_allclose(r.T @ np.diag(s ** 2) @ r, Gs[q], atol=g * B.TOL_BUDGET * ref[0] ** 2,
                                   err_msg=str(q))


def test_jacobi_gram_blocked_batch_mixed_orders():
    """One batched call over Gram matrices of mixed orders (partial last
    blocks, odd block counts, a graded spectrum, a rank-deficient one):
    every problem matches numpy's singular values (to the Gram route's
    eps s_max^2 / s_i) and reconstructs its G."""
    from syngular_amd import batch as B

    case("jacobi gram blocked batch, orders 65..300")
    rng = np.random.default_rng(11)
    orders = [65, 80, 96, 128, 129, 200, 256, 300, 128, 256]
    Cs, Gs = [], []
    for q, g in enumerate(orders):
        if q == 5:      # graded columns: singular values over 1e-3 .. 1
            C = rng.standard_normal((g + 5, g)) * np.logspace(0, -3, g)
        elif q == 7:    # rank 150 of 300
            C = rng.standard_normal((g, 150)) @ rng.standard_normal((150, g))
        else:
            C = rng.standard_normal((g + 3, g))
        Cs.append(C)
        Gs.append(C.T @ C)
    dev = torch.device("cuda")
    G = [torch.tensor(x, device=dev) for x in Gs]
    R = [torch.empty_like(x) for x in G]
    S = [torch.empty(x.shape[0], dtype=torch.float64, device=dev) for x in G]
    ptr = lambda ts: np.array([t.data_ptr() for t in ts], dtype=np.uint64)
    sweeps = B.jacobi_gram_blocked_table(ptr(G), np.array(orders, np.int64), ptr(R), ptr(S))
    torch.cuda.synchronize()
    assert 0 < sweeps < 60
    for q, g in enumerate(orders):
        s, r = S[q].cpu().numpy(), R[q].cpu().numpy()
        ref = np.linalg.svd(Cs[q], compute_uv=False)
        k = 150 if q == 7 else g
        assert np.all(np.diff(s) <= 0), q
        np.testing.assert_allclose(s[:k], ref[:k], rtol=1e-9, atol=1e-12 * ref[0], err_msg=str(q))
        np.testing.assert_allclose(r @ r.T, np.eye(g), atol=1e-12, err_msg=str(q))
        np.testing.assert_allclose(r.T @ np.diag(s ** 2) @ r, Gs[q], atol=1e-11 * ref[0] ** 2, err_msg=str(q))


@pytest.mark.parametrize("g", [511, 1024])
def test_jacobi_gram_blocked_graded(g):
    """Block Jacobi (16-row blocks: cross-block steps after the first of
    each pass) on a Gram matrix of a graded spectrum over four decades:
    singular values against LAPACK to the Gram route's accuracy (an error
    of about eps s_max^2 / s_i, at most ~2e-12 s_max here; the test allows
    5e-12 s_max), eigenvector rows orthonormal, the reconstruction within
    rounding."""
    from syngular_amd import ops

    case(f"jacobi gram blocked graded g={g}")
    rng = np.random.default_rng(g)
    C = rng.standard_normal((2 * g, g)) * np.logspace(0, -4, g)
    G0 = C.T @ C
    rt, s, sweeps = ops.jacobi_gram_blocked(T(G0))
    s = s.cpu().numpy()
    R = rt.cpu().numpy()
    ref = np.linalg.svd(C, compute_uv=False)
    assert np.all(np.diff(s) <= 0) and 0 < sweeps < 60
    np.testing.assert_allclose(s, ref, rtol=0, atol=5e-12 * ref[0])
    # value by value: the rotations' rounding leaves every value within about
    # g eps relative (7e-14 measured at the top of an order-511 spectrum), and
    # an eigenvalue of G is exact to about eps s_max^2 only, so s_i within a
    # further eps (s_max / s_i)^2 / 2 (1e-8 at s_i = 1e-4 s_max)
    eps = np.finfo(np.float64).eps
    rel = np.abs(s - ref) / ref
    bound = g * eps + 8 * eps * (ref[0] / ref) ** 2
    assert np.all(rel <= bound), float((rel / bound).max())
    np.testing.assert_allclose(R @ R.T, np.eye(g), atol=1e-12)
    np.testing.assert_allclose(R.T @ np.diag(s ** 2) @ R, G0, atol=1e-11 * ref[0] ** 2)


@pytest.mark.parametrize("cond,expect_full", [(10.0, False), (1e6, True)])
def test_cholqr_refine_paths(cond, expect_full):
    """Second CholeskyQR pass: the first-order factor of I + E when the first
    pass is already near-orthonormal, the full Cholesky otherwise; both give
    Q orthonormal and A = Q R."""
    from syngular_amd import ops

    case(f"cholqr refine cond={cond}")
    rng = np.random.default_rng(31)
    m, n = 900, 300
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * np.logspace(0, -np.log10(cond), n)) @ V.T
    Q1, L1, _ = ops.cholqr(T(A))
    Q, L2, full = ops.cholqr_refine(Q1)
    assert full == expect_full
    Qh = Q.cpu().numpy()
    R = (L2.t().cpu().numpy() @ L1.t().cpu().numpy())
    np.testing.assert_allclose(Qh.T @ Qh, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(Qh @ R, A, atol=1e-13)
    Q2, R2 = ops.cholqr2(T(A))
    np.testing.assert_allclose(Q2.cpu().numpy() @ R2.cpu().numpy(), A, atol=1e-13)


def test_orthogonality_estimate():
    """The randomised ||Q^T Q - I||_F estimate against the exact value, for a
    Q with a planted loss of orthogonality Q = U (I + F/2), F symmetric (so
    Q^T Q - I = F + F^2/4)."""
    from syngular_amd import ops

    case("orthogonality estimate")
    rng = np.random.default_rng(5)
    m, n = 1200, 400
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    for scale in (1e-12, 1e-9, 1e-6):
        F = rng.standard_normal((n, n)) * scale
        F = (F + F.T) / 2
        Q = U @ (np.eye(n) + F / 2)
        exact = np.linalg.norm(Q.T @ Q - np.eye(n))
        est = ops.orthogonality_estimate(T(Q))
        assert exact / 3 < est < 3 * exact, (scale, exact, est)
        assert est == ops.orthogonality_estimate(T(Q))   # fixed probes: deterministic


@pytest.mark.parametrize("cond,expect_skip", [(10.0, True), (1e6, False)])
def test_cholqr2_adaptive(cond, expect_skip):
    """Adaptive CholeskyQR: a well-conditioned matrix keeps its first pass
    (within the allowance), an ill-conditioned one runs the second; A = Q R
    both ways and Q^T Q - I stays within what was spent."""
    from syngular_amd import ops

    case(f"cholqr2 adaptive cond={cond}")
    rng = np.random.default_rng(32)
    m, n = 900, 300
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * np.logspace(0, -np.log10(cond), n)) @ V.T
    Q, R, spent = ops.cholqr2_adaptive(T(A), 1e-10)
    Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
    assert (spent > 0.0) == expect_skip
    np.testing.assert_allclose(Qh @ Rh, A, atol=1e-13)
    assert np.allclose(np.triu(Rh), Rh)
    E = np.linalg.norm(Qh.T @ Qh - np.eye(n))
    assert E <= (3 * spent if expect_skip else 1e-13 * n)


def test_round_budgeted_left_sweep():
    """TT-rounding with the budgeted left sweep (its default) against both
    CholeskyQR passes at every site: retained singular values within 1e-10
    relative, and against the oracle."""
    from syngular_amd import rounding
    from syngular_amd import workloads as W

    case("rounding budgeted left sweep N=12 chi=64 D_w=8")
    xs, os_ = W.chain_inputs(12, 2, 64, 8, seed=11)
    X = [T(x) for x in xs]
    O = [T(o) for o in os_]
    from syngular_amd import ops
    prod = ops.apply_many(list(zip(O, X)))
    r1, r2 = rounding.RoundingReport(), rounding.RoundingReport()
    rounding.right_svd_sweep(rounding.left_orthonormalize(prod, reorth_budget=rounding.ROUND_REORTH_BUDGET), 64,
                             0.0, r1)
    rounding.right_svd_sweep(rounding.left_orthonormalize(prod, reorth_budget=None), 64, 0.0, r2)
    ref, sv, kept, _ = ornd.apply_and_round(os_, xs, chi_max=64)
    assert r1.kept == r2.kept == kept[1:]
    for k in range(11):
        a, b = r1.svals[k].cpu().numpy(), r2.svals[k].cpu().numpy()
        c = kept[k + 1]
        assert np.max(np.abs(a[:c] - b[:c])) <= 1e-10 * b[0]
        assert np.max(np.abs(a[:c] - sv[k + 1][:c])) <= 1e-10 * sv[k + 1][0]


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(37, 50, 19), (300, 260, 513)])
def test_gemm_mixed(ta, tb, m, n, k):
    """Mixed GEMM: FP64 in/out, FP32 operands and accumulation; checks the
    FP32 MFMA output layout (an error would be O(1)) and the FP32 accuracy."""
    from syngular_amd import ops

    case(f"gemm mixed ta={ta} tb={tb} {m}x{n}x{k}")
    rng = np.random.default_rng(m + n + k + 3 * ta + tb)
    A = rng.standard_normal((k, m) if ta else (m, k))
    Bm = rng.standard_normal((n, k) if tb else (k, n))
    Ad, Bd = T(A), T(Bm)
    C = ops.gemm(Ad.t() if ta else Ad, Bd.t() if tb else Bd, mixed=True).cpu().numpy()
    opA, opB = (A.T if ta else A), (Bm.T if tb else Bm)
    ref = opA.astype(np.float32).astype(np.float64) @ opB.astype(np.float32).astype(np.float64)
    scale = np.abs(opA) @ np.abs(opB)
    assert np.abs(C - ref).max() / scale.max() < 4e-6
    assert np.abs(C - opA @ opB).max() / scale.max() < 1e-5


def test_mixed_round_vs_oracle():
    """The mixed form of apply+round on a chain against the FP64 oracle."""
    from oracle import reference as oref
    from oracle import rounding as ornd
    from syngular_amd import linalg, rounding
    from syngular_amd import workloads as W

    case("mixed apply+round chain")
    xs, os_ = W.chain_inputs(12, 2, 64, 8, seed=11)
    X = [T(x) for x in xs]
    O = [T(o) for o in os_]
    cores = rounding.apply_and_round(O, X, chi_max=64, precision="mixed")
    ref = ornd.apply_and_round(os_, xs, chi_max=64)[0]
    n2, n2r = float(linalg.overlap(cores, cores).item()), ornd.norm2(ref)
    ov, ovr = float(linalg.overlap(cores, X).item()), oref.overlap(ref, xs)
    # FP32 products move the state by ~2e-7 of its norm: the norm and the
    # overlap on the Cauchy-Schwarz scale ||Z|| ||X|| stay within 1e-6 of the
    # FP64 oracle; |<Z|X>| is only 0.12 of that scale on this chain, so the
    # overlap relative to itself is ~2e-6 (bounded here at 1e-5)
    assert abs(n2 - n2r) / n2r < 1e-6
    assert abs(ov - ovr) / np.sqrt(n2r * ornd.norm2(xs)) < 1e-6
    assert abs(ov - ovr) / abs(ovr) < 1e-5


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(37, 50, 19), (300, 260, 513), (1024, 512, 2048)])
def test_gemm_bf16(ta, tb, m, n, k):
    """BF16 GEMM: FP64 in/out, operands rounded to bfloat16, FP32
    accumulation; checks the bf16 MFMA operand and output layout (an error
    would be O(1)) against the product of the rounded operands, and the
    stated accuracy against the exact product."""
    from syngular_amd import ops

    case(f"gemm bf16 ta={ta} tb={tb} {m}x{n}x{k}")
    rng = np.random.default_rng(m + n + k + 5 * ta + tb)
    A = rng.standard_normal((k, m) if ta else (m, k))
    Bm = rng.standard_normal((n, k) if tb else (k, n))
    C = ops.gemm(T(A).t() if ta else T(A), T(Bm).t() if tb else T(Bm), mixed="bf16").cpu().numpy()
    opA, opB = (A.T if ta else A), (Bm.T if tb else Bm)

    def to_bf16(x):
        u = np.ascontiguousarray(x.astype(np.float32)).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
        return u.astype(np.uint32).view(np.float32).astype(np.float64)

    ref = to_bf16(opA) @ to_bf16(opB)
    scale = np.abs(opA) @ np.abs(opB)
    assert np.abs(C - ref).max() / scale.max() < 1e-5       # FP32 accumulation of exact bf16 products
    assert np.abs(C - opA @ opB).max() / scale.max() < 8e-3  # two bf16 roundings of 2^-9 each


def test_bf16_round_vs_oracle_states_its_error():
    """The bf16 form of apply+round (contractions on the BF16 matrix cores,
    QR products on the FP32 ones, Gram/Cholesky/Jacobi/truncation FP64) on a
    chain against the FP64 oracle: kept ranks equal, norm and overlap within
    the stated 1e-2 (BF16 keeps 8 bits of mantissa)."""
    from oracle import rounding as ornd
    from syngular_amd import linalg, rounding
    from syngular_amd import workloads as W

    case("bf16 apply+round chain")
    xs, os_ = W.chain_inputs(12, 2, 64, 8, seed=11)
    X = [T(x) for x in xs]
    O = [T(o) for o in os_]
    rep = rounding.RoundingReport()
    cores = rounding.apply_and_round(O, X, chi_max=64, report=rep, precision="bf16")
    ref, sv, kept, _ = ornd.apply_and_round(os_, xs, chi_max=64)
    assert rep.kept == kept[1:]
    n2, n2r = float(linalg.overlap(cores, cores).item()), ornd.norm2(ref)
    assert abs(n2 - n2r) / n2r < 1e-2
    ov = float(linalg.overlap(cores, X).item())
    from oracle import reference as oref
    assert abs(ov - oref.overlap(ref, xs)) / np.sqrt(n2r * ornd.norm2(xs)) < 1e-2
