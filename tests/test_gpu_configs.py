"""Parity at the benchmark configurations' own shapes (plain FP64 against
the NumPy FP64 oracle, relative 1e-10 per quantity unless a test says
otherwise).

cfg1 and cfg2 are long chains whose interior unfoldings only appear deep
inside; three-site open-boundary pieces reach exactly those shapes (the
8192 x 4096 CholeskyQR and 512-wide block Jacobi of cfg1, the 32768 x 8192
CholeskyQR and 4096-wide block Jacobi of cfg2) and are compared with oracle
fixtures (tests/golden/, written by make_config_fixtures.py).  The full
cfg1 chain is checked through size-independent identities.  cfg4 runs
served circuits at chi_max 64..512 against fixtures and alone against in a
mix; cfg3 runs a sample of its jobs against the live oracle.  Singular
values are compared value by value (|ds_i| / s_i), not relative to s_max.
"""
import os

import numpy as np
import pytest
import torch

from oracle import reference as oref
from oracle import rounding as ornd
from syngular_amd import batch as B
from syngular_amd import linalg, ops, rounding, serving
from syngular_amd import workloads as W

from golden import config_cases as cc

pytestmark = pytest.mark.gpu
TOL = 1e-10
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def case(name):
    print(f"[gpu case] {name}", flush=True)


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(ops.device())


def _fixture(name):
    path = os.path.join(GOLDEN, name)
    assert os.path.exists(path), f"{path} missing: run tests/golden/make_config_fixtures.py"
    return np.load(path)


def _per_value(got, want, tol=TOL):
    got, want = np.asarray(got), np.asarray(want)
    err = np.abs(got - want) / want
    assert err.max() <= tol, f"max per-value relative error {err.max():.3e} at index {int(err.argmax())}"
    return float(err.max())


def _orth_bound(cores):
    """Orthonormality a right core's rows keep: they are the Gram matrix's
    eigenvectors, accumulated over the Jacobi sweeps (up to ~8 here) from
    rotations of a few ulps each, so about 8 g eps for a g-wide Gram."""
    g = max(c.shape[1] * c.shape[2] for c in cores[1:])
    return 8 * g * np.finfo(np.float64).eps


def _check_interior(cores, svals, kept, disc, fx, label):
    """cores: host cores of the rounded open-boundary chain; svals/kept/disc
    per interior bond (1, 2)."""
    assert list(kept) == list(fx["kept"]), (label, kept, fx["kept"])
    for b, s in enumerate(svals):
        ref = fx[f"svals_{b + 1}"]
        c = int(fx["kept"][b])
        e = _per_value(s[:c], ref[:c])
        print(f"  {label} bond {b + 1}: {c} kept, per-value rel err {e:.2e}, s_min/s_max {ref[c - 1] / ref[0]:.2e}",
              flush=True)
        total = float(np.sum(ref ** 2))
        assert abs(disc[b] - fx["disc"][b]) <= TOL * total
    n2 = float(np.sum(cores[0] ** 2))
    assert abs(n2 - float(fx["norm2"])) <= TOL * float(fx["norm2"])
    us, ws = cc.boundary_vectors(cores[0].shape[0], cores[-1].shape[2])
    pr = cc.boundary_probes(cores, us, ws)
    for p, q in zip(pr, fx["probes"]):
        assert np.linalg.norm(p - q) <= TOL * np.linalg.norm(q)
    assert cc.right_orthonormality_error(cores) <= _orth_bound(cores)


def _interior_case(case_, fxname, label):
    xs, os_ = cc.case_inputs(case_)
    fx = _fixture(fxname)
    # single-chain path (CholeskyQR2 with split-K Gram, block Jacobi, vector norms)
    rep = rounding.RoundingReport()
    Z = rounding.apply_and_round([T(o) for o in os_], [T(x) for x in xs], chi_max=case_["chi_max"],
                                 cutoff=case_["cutoff"], report=rep)
    torch.cuda.synchronize()
    _check_interior([z.cpu().numpy() for z in Z], [s.cpu().numpy() for s in rep.svals], rep.kept,
                    [float(d.item()) for d in rep.discarded], fx, f"{label} chain")
    del Z, rep
    # grouped engine on the same piece
    X = B.ChainBatch.from_numpy([xs])
    O = B.OperatorBatch.from_numpy([os_])
    recs = []
    Zb = B.apply_and_round(X, O, case_["chi_max"], case_["cutoff"], records=recs)
    svals, kept, disc = [None, None], [None, None], [None, None]
    for r in recs:
        b = int(r.bond[0]) - 1
        g = int(r.g[0])
        out = torch.empty(g, dtype=torch.float64, device=ops.device())
        B.copy_table(r.s_ptr[:1], np.array([out.data_ptr()], np.uint64), np.array([g], np.int64))
        svals[b], kept[b], disc[b] = out.cpu().numpy(), int(r.kept[0]), float(r.disc[0].item())
    _check_interior(Zb.to_numpy()[0], svals, kept, disc, fx, f"{label} batch")


def test_cfg1_interior_sub_chain_vs_oracle():
    """cfg1 interior: chi=256, D_w=16, d=2; product bond 4096; two 8192 x
    4096 CholeskyQR sites and two 4096 x 512 right-sweep unfoldings."""
    case("cfg1 interior sub-chain (256, 2, 256) x 3, D_w=16, round to 256")
    _interior_case(cc.CFG1, "cfg1_interior.npz", "cfg1")


def test_cfg2_interior_sub_chain_vs_oracle():
    """cfg2 interior: chi=1024, D_w=8, d=4; product bond 8192; two 32768 x
    8192 CholeskyQR sites and two 8192 x 4096 right-sweep unfoldings;
    cutoff 1e-8 (the cap 1024 binds)."""
    case("cfg2 interior sub-chain (1024, 4, 1024) x 3, D_w=8, round to 1024, cutoff 1e-8")
    _interior_case(cc.CFG2, "cfg2_interior.npz", "cfg2")


def test_cfg1_full_chain_properties():
    """The whole cfg1 chain (N=32, d=2, chi=256, D_w=16): kept ranks
    min(256, 2^k, 2^(N-k)), Pythagoras ||OX||^2 - ||Z||^2 = sum_k w_k,
    right-orthonormal cores 1..N-1, sorted positive singular values."""
    case("cfg1 full chain N=32 d=2 chi=256 D_w=16")
    n = 32
    xs, os_ = W.chain_inputs(n, 2, 256, 16, seed=2024)
    X, O = [T(x) for x in xs], [T(o) for o in os_]
    rep = rounding.RoundingReport()
    Z = rounding.apply_and_round(O, X, chi_max=256, cutoff=0.0, report=rep)
    assert rep.kept == [min(256, 2 ** k, 2 ** (n - k)) for k in range(1, n)]
    P = ops.apply_many(list(zip(O, X)))
    full2 = float(linalg.overlap(P, P).item())
    del P
    z2 = float(linalg.overlap(Z, Z).item())
    wsum = sum(float(d.item()) for d in rep.discarded)
    assert abs(full2 - z2 - wsum) <= TOL * full2
    hz = [z.cpu().numpy() for z in Z]
    assert cc.right_orthonormality_error(hz) <= _orth_bound(hz)
    assert abs(float(np.sum(hz[0] ** 2)) - z2) <= TOL * z2
    for s, c in zip(rep.svals, rep.kept):
        v = s.cpu().numpy()[:c]
        assert (v > 0).all() and (np.diff(v) <= 1e-12 * v[0]).all()


def test_cfg2_reduced_chain_vs_oracle():
    """cfg2's d=4, D_w=8 and cutoff 1e-8 on a chain the oracle runs live
    (N=8, chi=64): per-bond singular values value by value, kept ranks,
    discarded weights, dense tensor, norm and overlap."""
    case("cfg2 reduced chain N=8 d=4 chi=64 D_w=8 cutoff 1e-8")
    n, d, chi, dw, cut = 8, 4, 64, 8, 1e-8
    xs, os_ = W.chain_inputs(n, d, chi, dw, seed=65)
    X, O = [T(x) for x in xs], [T(o) for o in os_]
    rep = rounding.RoundingReport()
    Z = rounding.apply_and_round(O, X, chi_max=chi, cutoff=cut, report=rep)
    ref, sv, kept, disc = ornd.apply_and_round(os_, xs, chi_max=chi, cutoff=cut)
    assert rep.kept == kept[1:]
    for k in range(1, n):
        _per_value(rep.svals[k - 1].cpu().numpy()[: kept[k]], sv[k][: kept[k]])
        assert abs(float(rep.discarded[k - 1].item()) - disc[k]) <= TOL * float(np.sum(sv[k] ** 2))
    hz = [z.cpu().numpy() for z in Z]
    a, b = oref.dense_mps(hz).reshape(-1), oref.dense_mps(ref).reshape(-1)
    assert np.linalg.norm(a - b) <= TOL * np.linalg.norm(b)
    n2r = ornd.norm2(ref)
    assert abs(float(linalg.overlap(Z, Z).item()) - n2r) <= TOL * n2r
    ov, ovr = float(linalg.overlap(Z, X).item()), oref.overlap(ref, xs)
    assert abs(ov - ovr) <= TOL * abs(ovr)


def _schmidt_check(got, want, layers):
    """Schmidt values of circuits rounded layer after layer: value by value
    within 1e-10, plus an absolute layers * 1e-12 * s_max.  Each truncating
    layer leaves the two computations' states apart by about 1e-13..1e-12 of
    the norm (rounding errors, amplified where a kept and a dropped singular
    value lie close), and a small retained Schmidt value sees that absolute
    difference (Weyl: |ds_i| <= ||d psi||)."""
    assert len(got) == len(want)
    err = np.abs(got - want) / (TOL * want + layers * 1e-12 * want[0])
    assert err.max() <= 1.0, f"Schmidt value {int(err.argmax())}: {got[int(err.argmax())]} vs {want[int(err.argmax())]}"


def test_cfg4_served_circuits_vs_oracle_and_alone():
    """cfg4's served form at chi_max 64, 128, 256 and 512 (cutoff 1e-10):
    four circuits from |0...0> through the continuous-batching server with
    lanes, against oracle fixtures (squared norm, discarded weight, bond
    dimensions, middle-bond Schmidt values, overlaps with product states),
    and each job alone against the same job in the mix."""
    case("cfg4 served circuits chi_max 64/128/256/512, cutoff 1e-10")
    jobs = cc.cfg4_jobs()
    fx = _fixture("cfg4_circuits.npz")
    srv = serving.Server(serving.DeviceExecutor.lanes(), hbm_budget=64 * 2 ** 30, keep_states=True,
                         lanes=serving.chi_lanes)
    rep = srv.run(jobs)
    torch.cuda.synchronize()
    mix = {r.id: r for r in rep.results}
    for j in jobs:
        q = j["id"]
        r = mix[q]
        n2r = float(fx[f"norm2_{q}"])
        assert abs(r.norm2 - n2r) <= TOL * n2r
        assert abs(r.discarded - float(fx[f"disc_{q}"])) <= TOL
        assert [c.shape[2] for c in r.cores[:-1]] == list(fx[f"bonds_{q}"])
        _schmidt_check(cc.schmidt_values(r.cores, j["n"] // 2), fx[f"schmidt_{q}"], j["layers"])
        vecs = cc.product_probes(j["n"])
        ov = cc.product_overlaps(r.cores, vecs)
        scale = np.sqrt(n2r) * np.prod(np.linalg.norm(vecs, axis=2), axis=1)
        assert (np.abs(ov - fx[f"overlaps_{q}"]) <= TOL * scale).all()
        print(f"  job {q} chi_max {j['chi_max']}: max bond {max(fx[f'bonds_{q}'])}, norm2 err "
              f"{abs(r.norm2 - n2r) / n2r:.1e}", flush=True)
    for j in jobs:
        solo = serving.Server(serving.DeviceExecutor(), hbm_budget=64 * 2 ** 30, keep_states=True).run([j])
        a, m = solo.results[0].cores, mix[j["id"]].cores
        assert [c.shape for c in a] == [c.shape for c in m]
        # the grouped GEMM's tile size follows the whole launch, so a job alone
        # and in a mix may round differently and near-degenerate singular
        # vectors may rotate: compare the states, not the gauge
        n2a, n2m = ornd.norm2(a), ornd.norm2(m)
        assert abs(n2a - n2m) <= TOL * n2m
        vecs = cc.product_probes(j["n"], seed=21)
        scale = np.sqrt(n2m) * np.prod(np.linalg.norm(vecs, axis=2), axis=1)
        assert (np.abs(cc.product_overlaps(a, vecs) - cc.product_overlaps(m, vecs)) <= TOL * scale).all()
        _schmidt_check(cc.schmidt_values(a, j["n"] // 2), cc.schmidt_values(m, j["n"] // 2), j["layers"])


def test_cfg3_batch_sample_vs_oracle():
    """cfg3's grouped form on 24 of its jobs (N ~ U[8,48], chi ~
    U{8..128}, one layer of Haar gates, rounded back to chi) against the
    live oracle: squared norm, overlap with the input and middle-bond
    Schmidt values value by value."""
    case("cfg3 batch sample of 24 jobs")
    jobs = W.batching_jobs(24, seed=33)
    X = B.ChainBatch.from_numpy([j["cores"] for j in jobs])
    Op = B.gate_layer([j["n"] for j in jobs], np.zeros(len(jobs), np.int64),
                      np.concatenate([j["gates"] for j in jobs]))
    Z = B.apply_and_round(X, Op, np.array([j["chi"] for j in jobs]))
    out = Z.to_numpy()
    for j, g in zip(jobs, out):
        ref = ornd.apply_and_round(W.gate_mpo_cores(j["n"], 0, j["gates"]), j["cores"], chi_max=j["chi"])[0]
        n2r = ornd.norm2(ref)
        assert abs(ornd.norm2(g) - n2r) <= TOL * n2r
        ovr = oref.overlap(ref, j["cores"])
        assert abs(oref.overlap(g, j["cores"]) - ovr) <= TOL * np.sqrt(n2r * ornd.norm2(j["cores"]))
        sg, sr = cc.schmidt_values(g, j["n"] // 2), cc.schmidt_values(ref, j["n"] // 2)
        assert len(sg) == len(sr)
        _per_value(sg, sr)


@pytest.mark.parametrize("rows,d,cols", [(400, 2, 100), (100, 2, 200), (96, 2, 24)])
def test_graded_spectrum_per_value(rows, d, cols):
    """Singular values of a right unfolding with a designed spectrum
    s_i = 10^(-5 i / (g - 1)) (down to 1e-5 s_max): every retained value
    within 1e-10 relative of numpy's SVD of the same matrix and 1e-9 of the
    design, through the single-chain and the grouped sweep; tall (block
    Jacobi), wide and LDS-Jacobi cases."""
    case(f"graded spectrum core ({rows}, {d}, {cols})")
    rng = np.random.default_rng(rows + cols)
    g = min(rows, d * cols)
    s = 10.0 ** (-5.0 * np.arange(g) / (g - 1))
    U = np.linalg.qr(rng.standard_normal((rows, g)))[0]
    V = np.linalg.qr(rng.standard_normal((d * cols, g)))[0]
    C = (U * s) @ V.T
    p = rows + 7
    Q0 = np.linalg.qr(rng.standard_normal((p, rows)))[0]
    cores = [Q0.reshape(1, p, rows), C.reshape(rows, d, cols), np.eye(cols).reshape(cols, cols, 1)]
    ref, sv, kept, _ = ornd.tt_round(cores, chi_max=None, cutoff=0.0)
    rep = rounding.RoundingReport()
    rounding.tt_round([T(c) for c in cores], chi_max=None, cutoff=0.0, report=rep)
    got = rep.svals[0].cpu().numpy()[: kept[1]]
    e1 = _per_value(got, sv[1][: kept[1]])
    _per_value(got, s[: kept[1]], tol=1e-9)
    recs = []
    B.ChainBatch.from_numpy([cores]).tt_round(0, 0.0, records=recs)
    r = [x for x in recs if int(x.bond[0]) == 1][0]
    out = torch.empty(int(r.g[0]), dtype=torch.float64, device=ops.device())
    B.copy_table(r.s_ptr[:1], np.array([out.data_ptr()], np.uint64), np.array([int(r.g[0])], np.int64))
    e2 = _per_value(out.cpu().numpy()[: kept[1]], sv[1][: kept[1]])
    print(f"  per-value rel err: chain {e1:.2e}, batch {e2:.2e}", flush=True)
