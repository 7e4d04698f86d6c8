"""The grouped many-chain sweep run on the CPU through the NumPy stand-in of
the C ABI (tests/abi_emulator.py): item tables, leading dimensions, block
counts and buffer extents are checked on every call, and the results are
compared with the NumPy FP64 oracle.  No GPU; the real kernels are covered by
tests/test_gpu_batch.py."""
import numpy as np
import pytest

import abi_emulator
from oracle import reference as oref
from oracle import rounding as ornd
from syngular_amd import batch as B
from syngular_amd import serving
from syngular_amd import workloads as W

TOL = 1e-10


@pytest.fixture
def emu(monkeypatch):
    return abi_emulator.install(monkeypatch)


def _chain(rng, n, chi, d=2):
    b = [1] + [chi] * (n - 1) + [1]
    return [rng.standard_normal((b[k], d, b[k + 1])) / np.sqrt(chi) for k in range(n)]


def _mpo(rng, n, D, d=2):
    w = [1] + [D] * (n - 1) + [1]
    return [rng.standard_normal((w[k], d, d, w[k + 1])) / np.sqrt(D * d) for k in range(n)]


SPEC = [(3, 2, 2), (5, 4, 3), (7, 8, 4), (4, 16, 2), (9, 6, 4), (14, 40, 3), (2, 5, 4), (8, 12, 1), (1, 3, 2), (10, 100, 2)]


def test_emulated_batched_round_matches_oracle(emu):
    rng = np.random.default_rng(0)
    jobs = [(_chain(rng, n, c), _mpo(rng, n, D), c) for n, c, D in SPEC]
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    recs = []
    Z = B.apply_and_round(X, O, np.array([j[2] for j in jobs]), 0.0, records=recs)
    out = Z.to_numpy()
    for fn in ("syn_jacobi_gram_blocked_f64", "syn_qr_sized_f64", "syn_eye_f64", "syn_chol_leaf_f64", "syn_jacobi_gram_f64"):
        assert fn in emu.calls
    for q, (xs, os_, c) in enumerate(jobs):
        ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=c)
        assert [o.shape for o in out[q]] == [r.shape for r in ref]
        a, b = oref.dense_mps(out[q]), oref.dense_mps(ref)
        assert np.abs(a - b).max() <= TOL * np.abs(b).max()
    for r in recs:
        for i, q in enumerate(r.jobs):
            assert int(r.kept[i]) == ornd.apply_and_round(jobs[q][1], jobs[q][0], chi_max=jobs[q][2])[2][int(r.bond[i])]


def test_emulated_cutoff_and_no_cap(emu):
    rng = np.random.default_rng(2)
    jobs = [(_chain(rng, n, c), _mpo(rng, n, D), c) for n, c, D in SPEC[:5]]
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    cut = np.array([1e-3, 0.0, 1e-2, 1e-6, 0.0])
    Z = B.apply_and_round(X, O, np.zeros(5, np.int64), cut)
    out = Z.to_numpy()
    for q, (xs, os_, c) in enumerate(jobs):
        ref = ornd.apply_and_round(os_, xs, chi_max=None, cutoff=cut[q])[0]
        assert [o.shape for o in out[q]] == [r.shape for r in ref]
        a, b = oref.dense_mps(out[q]), oref.dense_mps(ref)
        assert np.abs(a - b).max() <= TOL * np.abs(b).max()


def test_emulated_overlap(emu):
    rng = np.random.default_rng(4)
    A = [_chain(rng, n, c) for n, c in [(3, 2), (6, 5), (4, 9), (1, 1)]]
    C = [_chain(rng, n, c) for n, c in [(3, 3), (6, 2), (4, 9), (1, 1)]]
    ov = B.overlap(B.ChainBatch.from_numpy(A), B.ChainBatch.from_numpy(C)).numpy()
    for q in range(4):
        assert abs(ov[q] - oref.overlap(A[q], C[q])) <= 1e-12 * np.sqrt(ornd.norm2(A[q]) * ornd.norm2(C[q]))
    with pytest.raises(ValueError, match="edge bonds"):
        lit = [[np.ones((2, 2, 1))]]
        B.overlap(B.ChainBatch.from_numpy(lit), B.ChainBatch.from_numpy(lit))


def test_emulated_gate_layers(emu):
    jobs = W.batching_jobs(10, 7, n_range=(2, 9), chis=(2, 4, 8, 16, 64))
    X = B.ChainBatch.from_numpy([j["cores"] for j in jobs])
    for off in (0, 1):
        gates = [W.haar_orthogonal(np.random.default_rng(q), W.npairs(j["n"], off)) for q, j in enumerate(jobs)]
        O = B.gate_layer([j["n"] for j in jobs], np.full(len(jobs), off), np.concatenate(gates))
        recs = []
        Z = B.apply_and_round(X, O, np.array([j["chi"] for j in jobs]), 0.0, records=recs)
        out = Z.to_numpy()
        n2 = B.overlap(Z, Z).numpy()
        n2x = B.overlap(X, X).numpy()
        disc = np.zeros(len(jobs))
        for r in recs:
            np.add.at(disc, r.jobs, r.disc.numpy())
        for q, j in enumerate(jobs):
            mpo = W.gate_mpo_cores(j["n"], off, gates[q])
            ref = ornd.apply_and_round(mpo, j["cores"], chi_max=j["chi"])[0]
            a, b = oref.dense_mps(out[q]), oref.dense_mps(ref)
            assert np.abs(a - b).max() <= TOL * np.abs(b).max()
            assert abs(n2[q] + disc[q] - n2x[q]) <= 1e-10 * n2x[q]


def test_emulated_server_alone_equals_mix(emu):
    jobs = W.serving_jobs(5, 1e6, 3, n_range=(5, 9), layer_range=(2, 4), chis=(4, 8), cutoff=1e-10)
    rep = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True).run(jobs)
    mix = {r.id: r for r in rep.results}
    for j in jobs:
        cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
        for l in range(j["layers"]):
            mpo = W.gate_mpo_cores(j["n"], l % 2, W.layer_gates(j["seed"], j["id"], l, j["n"]))
            cores = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"], cutoff=j["cutoff"])[0]
        got = mix[j["id"]]
        a, b = oref.dense_mps(got.cores), oref.dense_mps(cores)
        assert np.abs(a - b).max() <= TOL
        assert abs(got.norm2 + got.discarded - 1.0) <= 1e-10
    solo = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True).run(
        [dict(jobs[2], arrival=0.0)])
    a, b = oref.dense_mps(solo.results[0].cores), oref.dense_mps(mix[jobs[2]["id"]].cores)
    assert np.abs(a - b).max() <= 1e-12


def test_emulator_catches_short_buffer(emu):
    """The stand-in must reject an item that reaches past its buffer."""
    a, pa = B.arena([4])
    b, pb = B.arena([4])
    c, pc = B.arena([4])
    with pytest.raises(abi_emulator.BoundsError):
        B.gemm_table(pa, pb, pc, [2], 2, 200, 200, 2, 2)


def test_emulated_server_longest_job_leaves_first(emu):
    """After the longest chain retires, the resident tables are narrower than
    before; later layers must still line up."""
    jobs = [{"id": 0, "arrival": 0.0, "n": 9, "layers": 1, "chi_max": 4, "cutoff": 0.0, "seed": 1},
            {"id": 1, "arrival": 0.0, "n": 5, "layers": 3, "chi_max": 4, "cutoff": 0.0, "seed": 1},
            {"id": 2, "arrival": 0.0, "n": 6, "layers": 2, "chi_max": 8, "cutoff": 0.0, "seed": 1}]
    rep = serving.Server(serving.DeviceExecutor(), hbm_budget=10 ** 11, keep_states=True).run(jobs)
    for r, j in zip(rep.results, jobs):
        cores = [np.eye(2)[0].reshape(1, 2, 1) for _ in range(j["n"])]
        for l in range(j["layers"]):
            mpo = W.gate_mpo_cores(j["n"], l % 2, W.layer_gates(j["seed"], j["id"], l, j["n"]))
            cores = ornd.apply_and_round(mpo, cores, chi_max=j["chi_max"])[0]
        assert np.abs(oref.dense_mps(r.cores) - oref.dense_mps(cores)).max() <= TOL


def test_emulated_cholqr2_table(emu):
    """Batched CholeskyQR2 over uneven sizes (leaf, one split, nested splits,
    a rank-deficient matrix): Q orthonormal on the kept columns, Q R = A."""
    rng = np.random.default_rng(5)
    shapes = [(80, 40), (300, 130), (700, 200), (600, 257), (400, 96)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats[4][:, 50:] = mats[4][:, :46] @ rng.standard_normal((46, 46))   # rank 50
    dev = [B._h2d(torch_from(a)) for a in mats]
    m = np.array([s[0] for s in shapes])
    n = np.array([s[1] for s in shapes])
    hold = []
    Q, R = B.cholqr2_table(np.array([t.data_ptr() for t in dev], np.uint64), m, n, n, hold)
    for q, (mm, nn) in enumerate(shapes):
        Qh = emu.buf(Q[q], mm * nn).reshape(mm, nn)
        Rh = emu.buf(R[q], nn * nn).reshape(nn, nn)
        assert np.abs(np.tril(Rh, -1)).max() == 0.0
        assert np.abs(Qh @ Rh - mats[q]).max() <= 1e-10 * np.abs(mats[q]).max()
        keep = np.abs(Qh).sum(axis=0) > 0
        if q < 4:
            assert keep.all()
        else:
            assert keep.sum() == 50
        Qk = Qh[:, keep]
        assert np.abs(Qk.T @ Qk - np.eye(keep.sum())).max() <= 1e-12


def torch_from(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a))


def test_emulated_mixed_precision_round(emu):
    """The mixed form (FP32 products) stays within FP32 rounding of the FP64
    oracle; the FP64 Gram/Jacobi/truncation keep the kept ranks."""
    rng = np.random.default_rng(0)
    jobs = [(_chain(rng, n, c), _mpo(rng, n, D), c) for n, c, D in SPEC]
    X = B.ChainBatch.from_numpy([j[0] for j in jobs])
    O = B.OperatorBatch.from_numpy([j[1] for j in jobs])
    out = B.apply_and_round(X, O, np.array([j[2] for j in jobs]), 0.0, precision="mixed").to_numpy()
    assert "syn_gemm_mixed_f64" in emu.calls
    for q, (xs, os_, c) in enumerate(jobs):
        ref = ornd.apply_and_round(os_, xs, chi_max=c)[0]
        assert [o.shape for o in out[q]] == [r.shape for r in ref]
        a, b = oref.dense_mps(out[q]), oref.dense_mps(ref)
        assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max()
    with pytest.raises(ValueError):
        B.apply_and_round(X, O, 4, precision="bf16")


def test_emulated_gaussian_batch_layout_and_round(emu):
    """ChainBatch.gaussian (cores drawn in place, grouped by chi in the arena)
    has the tables of the batching workload, cores of variance 1/chi, and
    rounds like the same cores given from the host; the refined singular
    values (vector norms of U S / S V^T) match the oracle's value by value."""
    ns, chis, gates = W.batching_layout(12, seed=5, n_range=(3, 9), chis=(2, 4, 8))
    X = B.ChainBatch.gaussian(ns, chis, 2, seed=5)
    host = X.to_numpy()
    for j, cores in enumerate(host):
        c = int(chis[j])
        assert [x.shape for x in cores] == [(1 if k == 0 else c, 2, c if k < ns[j] - 1 else 1)
                                            for k in range(int(ns[j]))]
    big = B.ChainBatch.gaussian([40] * 4, [16, 16, 64, 64], 2, seed=1).to_numpy()
    for cores, c in zip(big, (16, 16, 64, 64)):
        v = np.concatenate([x.reshape(-1) for x in cores[1:-1]])
        assert abs(v.var() * c - 1.0) < 0.05
    Op = B.gate_layer(ns, np.zeros(len(ns), np.int64), gates)
    recs = []
    Z = B.apply_and_round(X, Op, chis, 0.0, records=recs).to_numpy()
    goff = np.concatenate([[0], np.cumsum([W.npairs(int(v), 0) for v in ns])])
    for j in range(len(ns)):
        os_ = W.gate_mpo_cores(int(ns[j]), 0, gates[goff[j]:goff[j + 1]])
        ref = ornd.apply_and_round(os_, host[j], chi_max=int(chis[j]))[0]
        a, b = oref.dense_mps(Z[j]), oref.dense_mps(ref)
        assert np.abs(a - b).max() <= TOL * np.abs(b).max()
    assert "syn_vecnorms_f64" in emu.calls


def test_autotune_launch_classes_and_default_rule():
    from syngular_amd import autotune

    small = (np.full(10, 64), np.full(10, 64), np.full(10, 32))
    big = (np.full(300, 256), np.full(300, 256), np.full(300, 512))
    odd = (np.full(300, 130), np.full(300, 130), np.full(300, 512))
    assert autotune.tile(*small, 0) == 0            # too little area for 256 tiles of 128
    assert autotune.tile(*big, 0) == 1
    assert autotune.tile(*odd, 0) == 0              # 128-padding wastes more than 15% over 64's
    kb = autotune.launch_class(*big, 0)
    assert kb != autotune.launch_class(*big, 1) != autotune.launch_class(*big, 2)
    assert autotune.decode(autotune.launch_class(*big, 2))["mixed"] == "bf16"
    try:
        autotune.TABLE[kb] = 0
        assert autotune.tile(*big, 0) == 0
        autotune.collect()
        autotune.tile(*small, 0)
        autotune.tile(*big, 0)
        assert len(autotune._seen) == 2
    finally:
        autotune.collect(False)
        autotune.TABLE.clear()


def test_autotune_table_save_load(tmp_path):
    from syngular_amd import autotune

    try:
        autotune.TABLE.clear()
        autotune.TABLE.update({123: 1, 456: 0})
        autotune.TIMINGS.update({123: {0: 0.5, 1: 0.25}})
        autotune.COUNTS.update({123: 7})
        p = tmp_path / "t.json"
        autotune.save(str(p))
        autotune.TABLE.clear()
        autotune.load(str(p))
        assert autotune.TABLE == {123: 1, 456: 0}
        assert autotune.TIMINGS[123] == {0: 0.5, 1: 0.25} and autotune.COUNTS[123] == 7
    finally:
        autotune.TABLE.clear()
        autotune.TIMINGS.clear()
        autotune.COUNTS.clear()


def test_jacobi_tolerance_reaches_the_blocked_jacobi(emu, monkeypatch):
    """The served executor's jacobi_tol reaches every blocked Jacobi call;
    without it FP64 rounding keeps TOL_BLOCKED and the mixed form takes the
    error budget's TOL_BUDGET."""
    seen = []
    orig = B.jacobi_gram_blocked_table

    def record(x, g, rt, s, tol=None):
        seen.append(tol)
        return orig(x, g, rt, s, tol)

    monkeypatch.setattr(B, "jacobi_gram_blocked_table", record)
    jobs = W.serving_jobs(2, 1e6, 3, n_range=(12, 12), layer_range=(8, 8), chis=(32,), cutoff=1e-10)
    serving.Server(serving.DeviceExecutor(jacobi_tol=1e-8), hbm_budget=10 ** 11).run(jobs)
    assert seen and all(t == 1e-8 for t in seen)
    for prec, want in (("fp64", B.TOL_BLOCKED), ("mixed", B.TOL_BUDGET)):
        seen.clear()
        serving.Server(serving.DeviceExecutor(precision=prec), hbm_budget=10 ** 11).run(jobs)
        assert seen and all(t == want for t in seen), prec
