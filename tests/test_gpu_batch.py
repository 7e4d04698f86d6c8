"""Grouped many-chain path on the GPU against the NumPy FP64 oracle.

Gauge-invariant comparisons (relative 1e-10 for the plain FP64 form):
retained singular values per bond, dense reconstruction, norms and
overlaps; plus batch invariance (a chain alone and in a mix of others) and
the served loop."""
import numpy as np
import pytest
import torch

from oracle import reference as oref
from oracle import rounding as ornd
from syngular_amd import batch as B
from syngular_amd import ops, serving
from syngular_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _random_chain(rng, n, chi, d=2):
    b = [1] + [chi] * (n - 1) + [1]
    return [rng.standard_normal((b[k], d, b[k + 1])) / np.sqrt(chi) for k in range(n)]


def _random_mpo(rng, n, D, d=2):
    w = [1] + [D] * (n - 1) + [1]
    return [rng.standard_normal((w[k], d, d, w[k + 1])) / np.sqrt(D * d) for k in range(n)]


def _mixed_jobs(seed=0):
    rng = np.random.default_rng(seed)
    spec = [(3, 2, 2), (5, 4, 3), (7, 8, 4), (4, 16, 2), (9, 6, 4), (14, 40, 3), (2, 5, 4), (8, 12, 1), (10, 100, 2)]
    jobs = []
    for n, chi, D in spec:
        jobs.append((_random_chain(rng, n, chi), _random_mpo(rng, n, D), chi))
    return jobs


def _check_against_oracle(gpu_cores, ref_cores, n):
    if n <= 14:
        a = oref.dense_mps(gpu_cores)
        b = oref.dense_mps(ref_cores)
        assert np.abs(a - b).max() <= TOL * np.abs(b).max()
    n2g, n2r = ornd.norm2(gpu_cores), ornd.norm2(ref_cores)
    assert abs(n2g - n2r) <= TOL * n2r


def test_batched_round_matches_oracle_per_chain():
    jobs = _mixed_jobs()
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    chi = np.array([j[2] for j in jobs])
    recs = []
    Z = B.apply_and_round(X, O, chi, 0.0, records=recs)
    out = Z.to_numpy()
    for q, (xs, os_, c) in enumerate(jobs):
        ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=c)
        _check_against_oracle(out[q], ref, len(xs))
        assert [o.shape for o in out[q]] == [r.shape for r in ref]
        # per-bond singular values from the records
        for r in recs:
            w = np.nonzero(r.jobs == q)[0]
            if len(w) == 0:
                continue
            i = int(w[0])
            k = int(r.bond[i])
            g = int(r.g[i])
            s = torch.empty(g, dtype=torch.float64, device=ops.device())
            B.copy_table(np.array([r.s_ptr[i]], np.uint64), np.array([s.data_ptr()], np.uint64), np.array([g]))
            s = s.cpu().numpy()
            assert int(r.kept[i]) == kept[k]
            kk = kept[k]
            assert np.abs(s[:kk] - sv[k][:kk]).max() <= TOL * sv[k][0]


def test_batched_overlap_matches_oracle():
    rng = np.random.default_rng(4)
    A = [_random_chain(rng, n, c) for n, c in [(3, 2), (6, 5), (4, 9), (1, 1)]]
    Bc = [_random_chain(rng, n, c) for n, c in [(3, 3), (6, 2), (4, 9), (1, 1)]]
    ov = B.overlap(B.ChainBatch.from_numpy(A), B.ChainBatch.from_numpy(Bc)).cpu().numpy()
    for q in range(4):
        ref = oref.overlap(A[q], Bc[q])
        scale = np.sqrt(ornd.norm2(A[q]) * ornd.norm2(Bc[q]))
        assert abs(ov[q] - ref) <= TOL * scale


def test_chain_alone_equals_chain_in_mix():
    jobs = _mixed_jobs(1)
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    chi = np.array([j[2] for j in jobs])
    mix = B.apply_and_round(X, O, chi).to_numpy()
    for q in (0, 5, 7):
        Xa = B.ChainBatch.from_numpy([jobs[q][0]])
        Oa = B.OperatorBatch.from_numpy([jobs[q][1]])
        alone = B.apply_and_round(Xa, Oa, np.array([jobs[q][2]])).to_numpy()[0]
        assert [a.shape for a in alone] == [m.shape for m in mix[q]]
        da, dm = oref.dense_mps(alone), oref.dense_mps(mix[q])
        assert np.abs(da - dm).max() <= 1e-12 * np.abs(da).max()


def test_gate_layer_batch_matches_oracle_and_pythagoras():
    jobs = W.batching_jobs(12, 7, n_range=(4, 10), chis=(2, 4, 8, 16))
    X = B.ChainBatch.from_numpy([j["cores"] for j in jobs])
    O = B.gate_layer([j["n"] for j in jobs], np.zeros(len(jobs), np.int64), np.concatenate([j["gates"] for j in jobs]))
    recs = []
    Z = B.apply_and_round(X, O, np.array([j["chi"] for j in jobs]), 0.0, records=recs)
    out = Z.to_numpy()
    n2 = B.overlap(Z, Z).cpu().numpy()
    n2x = B.overlap(X, X).cpu().numpy()
    disc = np.zeros(len(jobs))
    for r in recs:
        np.add.at(disc, r.jobs, r.disc.cpu().numpy())
    for q, j in enumerate(jobs):
        mpo = W.gate_mpo_cores(j["n"], 0, j["gates"])
        ref, sv, kept, dref = ornd.apply_and_round(mpo, j["cores"], chi_max=j["chi"])
        _check_against_oracle(out[q], ref, j["n"])
        # orthogonal gates keep the norm: ||Z||^2 + sum of discarded = ||X||^2
        assert abs(n2[q] + disc[q] - n2x[q]) <= 1e-10 * n2x[q]


def test_blocked_many_jacobi_matches_single():
    rng = np.random.default_rng(9)
    sizes = [70, 96, 130, 200]
    mats = []
    for g in sizes:
        C = rng.standard_normal((g + 20, g))
        mats.append(torch.from_numpy(C.T @ C).to(ops.device()))
    gt, gp = B.arena([g * g for g in sizes])
    B.copy_table(np.array([m.data_ptr() for m in mats], np.uint64), gp, np.array([g * g for g in sizes]))
    rtt, rtp = B.arena([g * g for g in sizes])
    st, sp = B.arena(sizes)
    B.jacobi_blocked_table(gp, np.array(sizes), rtp, sp)
    torch.cuda.synchronize()
    host = st.cpu().numpy()
    base = np.uint64(st.data_ptr())
    for q, g in enumerate(sizes):
        o = int((sp[q] - base) // np.uint64(8))
        s = host[o:o + g]
        ref = np.sqrt(np.linalg.eigvalsh(mats[q].cpu().numpy())[::-1].clip(min=0))
        assert np.abs(s - ref).max() <= 1e-12 * ref[0]


def test_served_job_alone_equals_in_mix():
    jobs = W.serving_jobs(6, 1e6, 3, n_range=(6, 12), layer_range=(2, 5), chis=(4, 8, 16), cutoff=1e-10)
    srv = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True)
    rep = srv.run(jobs)
    mix = {r.id: r for r in rep.results}
    for j in (jobs[1], jobs[4]):
        solo = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True).run([dict(j, arrival=0.0)])
        a, m = solo.results[0], mix[j["id"]]
        assert [c.shape for c in a.cores] == [c.shape for c in m.cores]
        da, dm = oref.dense_mps(a.cores), oref.dense_mps(m.cores)
        assert np.abs(da - dm).max() <= 1e-12
    for r in rep.results:
        # unitary layers: norm^2 + discarded weight = 1
        assert abs(r.norm2 + r.discarded - 1.0) <= 1e-10
        assert r.latency >= 0 and r.layers == next(j["layers"] for j in jobs if j["id"] == r.id)


def test_served_job_matches_oracle_circuit():
    jobs = W.serving_jobs(3, 1e6, 5, n_range=(6, 9), layer_range=(3, 4), chis=(4, 8), cutoff=1e-10)
    rep = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True).run(jobs)
    for r, j in zip(rep.results, jobs):
        cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
        for l in range(j["layers"]):
            mpo = W.gate_mpo_cores(j["n"], l % 2, W.layer_gates(j["seed"], j["id"], l, j["n"]))
            cores = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])[0]
        _check_against_oracle(r.cores, cores, j["n"])


def test_gram_blocked_many_uneven():
    rng = np.random.default_rng(19)
    sizes = [65, 96, 130, 200, 33 * 16 + 5]
    mats = []
    for g in sizes:
        C = rng.standard_normal((g + 20, g))
        mats.append(torch.from_numpy(C.T @ C).to(ops.device()))
    gt, gp = B.arena([g * g for g in sizes])
    B.copy_table(np.array([m.data_ptr() for m in mats], np.uint64), gp, np.array([g * g for g in sizes]))
    rtt, rtp = B.arena([g * g for g in sizes])
    st, sp = B.arena(sizes)
    B.jacobi_gram_blocked_table(gp, np.array(sizes), rtp, sp)
    torch.cuda.synchronize()
    hs, hr = st.cpu().numpy(), rtt.cpu().numpy()
    for q, g in enumerate(sizes):
        o = int((sp[q] - np.uint64(st.data_ptr())) // np.uint64(8))
        orr = int((rtp[q] - np.uint64(rtt.data_ptr())) // np.uint64(8))
        s = hs[o:o + g]
        R = hr[orr:orr + g * g].reshape(g, g)
        G0 = mats[q].cpu().numpy()
        ref = np.sqrt(np.linalg.eigvalsh(G0)[::-1].clip(min=0))
        assert np.abs(s - ref).max() <= 1e-12 * ref[0]
        assert np.abs(R.T @ np.diag(s ** 2) @ R - G0).max() <= 1e-11 * ref[0] ** 2


def test_served_lanes_alone_equals_in_mix():
    """Lanes (one host thread and HIP stream each) give every job the same
    result as when it runs alone."""
    jobs = W.serving_jobs(6, 1e6, 8, n_range=(6, 12), layer_range=(2, 5), chis=(4, 8, 16), cutoff=1e-10)
    lanes = lambda j: 0 if j["chi_max"] <= 8 else 1
    rep = serving.Server(serving.DeviceExecutor.lanes(), hbm_budget=10 ** 11, keep_states=True,
                         lanes=lanes).run(jobs)
    torch.cuda.synchronize()
    mix = {r.id: r for r in rep.results}
    for j in jobs:
        solo = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True).run(
            [dict(j, arrival=0.0)])
        a, m = solo.results[0], mix[j["id"]]
        da, dm = oref.dense_mps(a.cores), oref.dense_mps(m.cores)
        assert np.abs(da - dm).max() <= 1e-12
        assert abs(m.norm2 + m.discarded - 1.0) <= 1e-10


def test_batched_orthogonality_budget():
    """The batched estimate of ||Q^T Q - I||_F against the exact value for
    planted losses, and the budgeted left sweep against both CholeskyQR
    passes everywhere: per-bond singular values within 1e-10 relative."""
    rng = np.random.default_rng(77)
    qs, exact = [], []
    for (m, n, scale) in ((300, 120, 1e-12), (500, 200, 1e-9), (260, 130, 1e-6)):
        U, _ = np.linalg.qr(rng.standard_normal((m, n)))
        F = rng.standard_normal((n, n)) * scale
        F = (F + F.T) / 2
        Q = U @ (np.eye(n) + F / 2)
        qs.append(torch.from_numpy(Q).to(ops.device()))
        exact.append(np.linalg.norm(Q.T @ Q - np.eye(n)))
    hold = []
    est = B.orth_estimate_table(np.array([q.data_ptr() for q in qs], np.uint64),
                                np.array([q.shape[0] for q in qs]), np.array([q.shape[1] for q in qs]), hold)
    torch.cuda.synchronize()
    for e, x in zip(est, exact):
        assert x / 3 < e < 3 * x, (e, x)

    jobs = _mixed_jobs(3)
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    chi = np.array([j[2] for j in jobs])
    P = X.apply(O)
    ra, rb = [], []
    A = P.left_orthonormalize(False, B.REORTH_BUDGET).right_svd_sweep(chi, 0.0, ra, False)
    Bb = P.left_orthonormalize(False, None).right_svd_sweep(chi, 0.0, rb, False)
    for q, (xs, os_, c) in enumerate(jobs):
        ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=c)
        _check_against_oracle(A.to_numpy()[q], ref, len(xs))
        n2a, n2b = ornd.norm2(A.to_numpy()[q]), ornd.norm2(Bb.to_numpy()[q])
        assert abs(n2a - n2b) <= 1e-10 * n2b


def test_autotuned_tiles_keep_results_and_batch_invariance():
    """The tile autotuner: record the launch classes of a mixed batch, time
    both tile configurations per class, then the tuned run still matches the
    oracle and a chain alone agrees with the same chain in the mix (to
    rounding: the tile follows the launch)."""
    from syngular_amd import autotune

    jobs = _mixed_jobs(5)
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    chi = np.array([j[2] for j in jobs])
    try:
        autotune.collect()
        B.apply_and_round(X, O, chi)
        table = autotune.tune()
        assert table and all(t in (0, 1) for t in table.values())
        out = B.apply_and_round(X, O, chi).to_numpy()
        for q, (xs, os_, c) in enumerate(jobs):
            ref = ornd.apply_and_round(os_, xs, chi_max=c)[0]
            _check_against_oracle(out[q], ref, len(xs))
        for q in range(len(jobs)):
            solo = B.apply_and_round(B.ChainBatch.from_numpy([jobs[q][0]]),
                                     B.OperatorBatch.from_numpy([jobs[q][1]]), np.array([jobs[q][2]])).to_numpy()[0]
            _check_against_oracle(solo, out[q], len(jobs[q][0]))
    finally:
        autotune.TABLE.clear()


def test_captured_apply_left_sweep_replays_match_eager():
    """The apply and left sweep of a batch captured in a HIP graph: four
    replays, with eager work of other forms and synchronisations between
    them, each giving the eager mixed-form result bit for bit."""
    jobs = _mixed_jobs(7)
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    chi = np.array([j[2] for j in jobs])
    eager = B.apply_and_round(X, O, chi, precision="mixed").to_numpy()
    cap = B.CapturedApplyLeft(X, O, mixed=True)
    assert cap.staging.tables > 0
    for rep in range(4):
        got = cap.round(chi).to_numpy()
        for q in range(len(jobs)):
            assert [c.shape for c in got[q]] == [c.shape for c in eager[q]]
            assert all(np.array_equal(a, b) for a, b in zip(got[q], eager[q])), (rep, q)
        B.apply_and_round(X, O, chi, precision="fp64")
        torch.cuda.synchronize()
    for q, (xs, os_, c) in enumerate(jobs):
        ref = ornd.apply_and_round(os_, xs, chi_max=c)[0]
        n2r = ornd.norm2(ref)
        assert abs(ornd.norm2(got[q]) - n2r) <= 1e-6 * n2r


@pytest.mark.parametrize("kind", [False, True, "bf16"])
def test_grouped_gemm_persistent_launch(kind):
    """A grouped launch of more than 4096 tiles runs persistent (a fixed grid
    walking the tiles): 5000 problems of uneven shape, each against NumPy
    (FP64 to 1e-13; the FP32 and BF16 kinds to their operand rounding)."""
    from syngular_amd import ops as O_

    rng = np.random.default_rng(71)
    cnt = 5000
    m = rng.integers(1, 65, cnt)
    n = rng.integers(1, 65, cnt)
    k = rng.integers(1, 90, cnt)
    A = [rng.standard_normal((a, c)) for a, c in zip(m, k)]
    Bm = [rng.standard_normal((c, b)) for c, b in zip(k, n)]
    ta, pa = B.arena([a.size for a in A])
    tb, pb = B.arena([b.size for b in Bm])
    tc, pc = B.arena(m * n)
    ha, hb = ta.cpu().numpy(), tb.cpu().numpy()
    base_a, base_b = np.uint64(ta.data_ptr()), np.uint64(tb.data_ptr())
    for i in range(cnt):
        oa, ob = int((pa[i] - base_a) // np.uint64(8)), int((pb[i] - base_b) // np.uint64(8))
        ha[oa:oa + A[i].size] = A[i].reshape(-1)
        hb[ob:ob + Bm[i].size] = Bm[i].reshape(-1)
    ta.copy_(torch.from_numpy(ha))
    tb.copy_(torch.from_numpy(hb))
    B.gemm_table(pa, pb, pc, m, n, k, k, n, n, mixed=kind)
    hc = tc.cpu().numpy()
    base_c = np.uint64(tc.data_ptr())
    tol = {False: 1e-13, True: 1e-5, "bf16": 8e-3}[kind]
    for i in range(cnt):
        oc = int((pc[i] - base_c) // np.uint64(8))
        got = hc[oc:oc + m[i] * n[i]].reshape(m[i], n[i])
        scale = np.abs(A[i]) @ np.abs(Bm[i])
        assert np.abs(got - A[i] @ Bm[i]).max() <= tol * scale.max(), i
