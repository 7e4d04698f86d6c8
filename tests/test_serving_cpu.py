"""Host-side logic of the many-chain path: the gate -> MPO convention, the
workload generators and the continuous-batching scheduler (driven by a
stand-in executor and a simulated clock; no GPU)."""
import numpy as np
import pytest

from oracle import reference as oref
from oracle import rounding as ornd
from syngular_amd import serving
from syngular_amd import workloads as W
from syngular_amd.batch import gate_pair_cores


def test_haar_orthogonal():
    q = W.haar_orthogonal(np.random.default_rng(1), 64)
    assert np.abs(q @ q.transpose(0, 2, 1) - np.eye(4)).max() < 1e-14
    # both determinant signs occur (Haar on O(4), not SO(4))
    assert set(np.sign(np.linalg.det(q)).astype(int)) == {-1, 1}


def test_gate_pair_cores_apply_U():
    """The two MPO cores of a gate, contracted and applied with the
    reference's MPO x MPS rule, give U x on the pair (dense check)."""
    rng = np.random.default_rng(3)
    U = rng.standard_normal((5, 4, 4))
    A, B = gate_pair_cores(U)
    assert A.shape == (5, 1, 2, 2, 4) and B.shape == (5, 4, 2, 2, 1)
    for p in range(5):
        W2 = oref.dense_mpo([A[p], B[p]])           # [in0, in1, out0, out1]
        assert np.allclose(W2.reshape(4, 4), U[p].T, atol=1e-13)
        x = [rng.standard_normal((1, 2, 3)), rng.standard_normal((3, 2, 1))]
        y = ornd.apply_mpo([A[p], B[p]], x)
        assert np.allclose(oref.dense_mps(y).reshape(4), U[p] @ oref.dense_mps(x).reshape(4), atol=1e-13)


def test_gate_mpo_cores_layer_layout():
    gates = W.haar_orthogonal(np.random.default_rng(0), 3)
    cores = W.gate_mpo_cores(7, 1, gates)
    assert [c.shape[0] for c in cores] == [1, 1, 4, 1, 4, 1, 4]
    assert [c.shape[3] for c in cores] == [1, 4, 1, 4, 1, 4, 1]
    dense = oref.dense_mpo(cores).reshape(2 ** 7, 2 ** 7)
    assert np.abs(dense @ dense.T - np.eye(2 ** 7)).max() < 1e-12


def test_layer_gates_keyed_by_job_and_layer():
    a = W.layer_gates(5, 17, 3, 20)
    b = W.layer_gates(5, 17, 3, 20)
    c = W.layer_gates(5, 18, 3, 20)
    assert a.shape == (9, 4, 4) and np.array_equal(a, b) and not np.allclose(a, c)
    assert W.layer_gates(5, 17, 2, 20).shape == (10, 4, 4)


def test_generators_ranges():
    jobs = W.batching_jobs(50, 0)
    assert all(8 <= j["n"] <= 48 and j["chi"] in (8, 16, 32, 64, 128) for j in jobs)
    j = jobs[0]
    assert j["cores"][0].shape == (1, 2, j["chi"]) and j["cores"][-1].shape == (j["chi"], 2, 1)
    assert j["gates"].shape == (j["n"] // 2, 4, 4)
    s = W.serving_jobs(200, 10.0, 1)
    arr = np.array([x["arrival"] for x in s])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    assert abs(np.mean(np.diff(arr)) - 0.1) < 0.03
    assert all(16 <= x["n"] <= 128 and 4 <= x["layers"] <= 32 and x["chi_max"] in (64, 128, 256, 512) for x in s)


# --------------------------------------------------------------------------
# scheduler with a stand-in executor and a simulated clock
# --------------------------------------------------------------------------
class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class FakeExecutor:
    """State = list of job ids; every layer costs `dt` simulated seconds."""

    def __init__(self, clock, dt=1.0):
        self.clock, self.dt = clock, dt
        self.calls = []

    def empty(self):
        return []

    def init(self, jobs):
        return [j["id"] for j in jobs]

    def merge(self, a, b):
        return a + b

    def select(self, state, idx):
        return [state[i] for i in idx]

    def layer(self, state, jobs, done):
        assert state == [j["id"] for j in jobs]
        self.calls.append(list(zip(state, done)))
        self.clock.t += self.dt
        return state, {"discarded": [0.01] * len(state)}

    def harvest(self, state, idx, keep_states=False):
        return {"norm2": [1.0 - 0.01 * 0] * len(idx)}


def _job(i, arrival, n=16, layers=3, chi=64):
    return {"id": i, "arrival": arrival, "n": n, "layers": layers, "chi_max": chi, "cutoff": 1e-10, "seed": 0}


def test_continuous_batching_joins_at_layer_boundary():
    clk = FakeClock()
    ex = FakeExecutor(clk)
    srv = serving.Server(ex, hbm_budget=10 ** 12, clock=clk, sleep=clk.sleep)
    jobs = [_job(0, 0.0, layers=4), _job(1, 1.5, layers=2), _job(2, 10.0, layers=1)]
    rep = srv.run(jobs)
    # layer 1 at t=0: job 0 alone; job 1 arrives at 1.5 and joins at t=2
    assert ex.calls[0] == [(0, 0)]
    assert ex.calls[1] == [(0, 1)]
    assert ex.calls[2] == [(0, 2), (1, 0)]
    assert ex.calls[3] == [(0, 3), (1, 1)]
    # idle until job 2 arrives at t=10
    assert ex.calls[4] == [(2, 0)]
    r = {x.id: x for x in rep.results}
    assert r[0].finished == 4.0 and r[1].finished == 4.0 and r[2].finished == 11.0
    assert r[1].latency == pytest.approx(2.5) and r[2].latency == pytest.approx(1.0)
    assert r[0].discarded == pytest.approx(0.04)
    assert rep.site_layers == 16 * (4 + 2 + 1)
    assert rep.admitted_during_run == 1
    assert rep.latency_percentiles()["p50"] == pytest.approx(2.5)


def test_admission_respects_hbm_budget_and_slots():
    clk = FakeClock()
    ex = FakeExecutor(clk)
    one = serving.job_bytes(16, 64)
    srv = serving.Server(ex, hbm_budget=2 * one + one // 2, clock=clk, sleep=clk.sleep)
    jobs = [_job(i, 0.0, layers=2) for i in range(5)]
    rep = srv.run(jobs)
    assert max(len(c) for c in ex.calls) == 2
    assert rep.peak_bytes <= 2 * one + one // 2
    assert sorted(x.finished for x in rep.results) == [2.0, 2.0, 4.0, 4.0, 6.0]
    clk2 = FakeClock()
    ex2 = FakeExecutor(clk2)
    rep2 = serving.Server(ex2, hbm_budget=10 ** 15, max_slots=3, clock=clk2, sleep=clk2.sleep).run(jobs)
    assert max(len(c) for c in ex2.calls) == 3 and rep2.peak_resident == 3


def test_job_larger_than_budget_is_rejected():
    clk = FakeClock()
    srv = serving.Server(FakeExecutor(clk), hbm_budget=1000, clock=clk, sleep=clk.sleep)
    with pytest.raises(ValueError, match="HBM budget"):
        srv.run([_job(0, 0.0)])


def test_job_bytes_formula():
    assert serving.job_bytes(100, 256) == serving.WORKING_SET_FACTOR * 8 * 100 * 2 * 256 * 256


class SleepExecutor:
    """Stand-in lane executor on the real clock: a layer costs `cost(job)`
    seconds for its heaviest job (lock-step within the lane)."""

    def __init__(self, log, lock):
        self.log, self.lock = log, lock

    def empty(self):
        return []

    def init(self, jobs):
        return [j["id"] for j in jobs]

    def merge(self, a, b):
        return a + b

    def select(self, state, idx):
        return [state[i] for i in idx]

    def layer(self, state, jobs, done):
        import time as _t
        _t.sleep(max(0.002 * (j["chi_max"] // 64) for j in jobs))
        with self.lock:
            self.log.extend((j["id"], l) for j, l in zip(jobs, done))
        return state, {"discarded": [0.0] * len(state)}

    def harvest(self, state, idx, keep_states=False):
        return {"norm2": [1.0] * len(idx)}


def test_lanes_decouple_light_jobs_from_heavy_ones():
    import threading
    log, lock = [], threading.Lock()
    jobs = [_job(0, 0.0, layers=10, chi=512)] + [_job(i, 0.0, layers=3, chi=64) for i in range(1, 6)]
    one = serving.job_bytes(16, 512)
    srv = serving.Server(lambda key: SleepExecutor(log, lock), hbm_budget=2 * one, lanes=serving.chi_lanes)
    rep = srv.run(jobs)
    r = {x.id: x for x in rep.results}
    # every job ran its layers in order
    for j in jobs:
        assert [l for i, l in log if i == j["id"]] == list(range(j["layers"]))
    # light jobs (3 layers of ~2 ms) do not wait for the heavy one (10 x ~16 ms)
    assert max(r[i].finished for i in range(1, 6)) < r[0].finished
    assert rep.peak_bytes <= 2 * one
    assert rep.site_layers == 16 * (10 + 5 * 3)
    assert serving.chi_lanes({"chi_max": 64}) == 0 and serving.chi_lanes({"chi_max": 512}) == 2


def test_lane_errors_propagate():
    class Boom(SleepExecutor):
        def layer(self, state, jobs, done):
            raise RuntimeError("boom")
    import threading
    srv = serving.Server(lambda key: Boom([], threading.Lock()), hbm_budget=10 ** 12, lanes=serving.chi_lanes)
    with pytest.raises(RuntimeError, match="boom"):
        srv.run([_job(0, 0.0, chi=64), _job(1, 0.0, chi=512)])
