"""Dense gates and the qubit register on the GPU against the oracle.

MPS.apply (tensor/matrix_product_state.py:487-534) and Qbit / Circuit
(quantum/qbit.py, quantum/circuit.py) through the device path, compared with
the NumPy oracle on the same inputs.  Dense tensors agree to a relative
Frobenius error <= 1e-10.  The split uses Householder reflectors with the
LAPACK convention, so for generic inputs the new cores themselves match the
oracle's (numpy.linalg.qr) to the same tolerance, not only their product.
"""
import numpy as np
import pytest

from oracle import quantum as QO
from oracle import reference as R

pytestmark = pytest.mark.gpu

TOL = 1e-10


def case(name):
    print(f"[gpu case] {name}", flush=True)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def random_sites(rng, dims, bonds):
    b = [1] + list(bonds) + [1]
    return [rng.standard_normal((b[k], d, b[k + 1])) for k, d in enumerate(dims)]


def sites_np(mp):
    return [s.cpu().numpy() for s in mp.sites]


@pytest.mark.parametrize("dims,bonds,m,index", [
    ((2, 2, 2), (2, 2), 2, 1),                 # test/test_mps.py:9-19 shape
    ((2, 3, 2, 4), (2, 5, 3), 1, 1),           # one-site gate, d = 3
    ((2, 3, 2, 3, 2), (2, 4, 3, 2), 2, 1),     # truncating split (rank 4 of 6)
    ((2, 2, 2, 2, 2, 2), (2, 4, 8, 4, 2), 2, 2),
    ((2, 2, 2), (2, 2), 3, 0),                 # three sites, r_1 = d_2 r_2
    ((4, 4, 4, 4), (4, 16, 4), 2, 1),
    ((2,) * 8, (2, 4, 8, 12, 8, 4, 2), 2, 3),  # truncating split (rank 12 of 16)
])
def test_apply_matches_oracle(dims, bonds, m, index):
    from syngular_amd.tensor import MatrixProductState as MPS

    case(f"apply dims={dims} bonds={bonds} m={m} index={index}")
    rng = np.random.default_rng(hash((dims, m, index)) % 2 ** 32)
    sites = random_sites(rng, dims, bonds)
    d = dims[index:index + m]
    G = rng.standard_normal(tuple(d) + tuple(d))
    want = R.MPS.from_sites(sites).apply(G, index)
    X = MPS.from_sites(sites)
    got = X.apply(G, index)
    assert got.shape == want.shape
    assert rel(got.to_tensor(), want.to_tensor()) < TOL
    for a, b in zip(sites_np(got), want.sites):
        assert rel(a, b) < TOL
    for a, b in zip(sites_np(X), sites):        # receiver untouched
        np.testing.assert_array_equal(a, b)


def test_apply_known_answer_and_errors():
    from syngular_amd.quantum import gate
    from syngular_amd.tensor import MatrixProductState as MPS

    case("apply CX arange(8) and error paths")
    x = np.arange(8.0).reshape(2, 2, 2)
    X = MPS(x, bond_shape=(2, 2)).decompose()
    want = x.copy()
    want[:, 1, :] = x[:, 1, ::-1]
    assert rel(X.apply(gate.CX, 1).to_tensor(), want) < TOL
    with pytest.raises(UnboundLocalError):
        X.apply(gate.CX, 0, strict=False)
    with pytest.raises(ValueError):
        X.apply(np.ones((2, 2, 2)), 0)
    with pytest.raises(TypeError):
        X.apply(gate.S, 0)
    # S @ S = Z up to a rounding-size imaginary part: still complex, rejected
    with pytest.raises(TypeError):
        X.apply(gate.S @ gate.S, 0)
    # exactly real complex arrays are accepted
    assert rel(X.apply(gate.Z.astype(complex), 0).to_tensor(), R.MPS(x, (2, 2)).decompose().apply(QO.Z, 0).to_tensor()) < TOL
    rng = np.random.default_rng(7)
    Y = MPS.from_sites(random_sites(rng, (2, 2, 2, 2), (2, 3, 2)))
    with pytest.raises(ValueError):
        Y.apply(rng.standard_normal((2,) * 6), 0)      # middle bond 3 != d r


def test_gates_match_reference_matrices():
    from syngular_amd.quantum import gate

    case("gate constants")
    for name in ("I", "X", "Z", "H", "CX", "SWAP", "TOFFOLI"):
        np.testing.assert_array_equal(getattr(gate, name), getattr(QO, name))


def _circuits():
    n = 4
    token = "1011"
    bv = [(QO.H, n), (QO.Z, n)] + [(QO.H, i) for i in range(n)]
    bv += [(QO.CX, n, i) for i in range(n) if token[i] == "1"] + [(QO.H, i) for i in range(n)]
    return {
        "bell": (2, [(QO.H, 0), (QO.CX, 0, 1)]),
        "ghz6": (6, [(QO.H, 0)] + [(QO.CX, i, i + 1) for i in range(5)]),
        "bernstein_vazirani_1011": (n + 1, bv),
        "swap_quirk": (3, [(QO.H, 0), (QO.X, 1), (QO.CX, 0, 2)]),
        "swap_far": (5, [(QO.X, 4), (QO.H, 0), (QO.CX, 4, 0), (QO.H, 2)]),
        "toffoli": (3, [(QO.X, 0), (QO.X, 1), (QO.TOFFOLI, 0)]),
        "staircase": (6, staircase(6)),
    }


def staircase(n, seed=5):
    """H, then random orthogonal two-qubit gates on (0,1), (1,2), ...  The
    split keeps the span of the first two columns of each unfolding, which
    here drops weight (the final norm is about 0.71), but the result is
    stable: perturbing every core by 1e-15 moves it by ~1e-14.  Deeper
    circuits are left out: there the kept span comes from rank-deficient
    unfoldings and rounding-size perturbations change the result by O(1),
    in the oracle itself."""
    rng = np.random.default_rng(seed)
    ops = [(QO.H, 0)]
    for i in range(n - 1):
        q, _ = np.linalg.qr(rng.standard_normal((4, 4)))
        ops.append((q.reshape(2, 2, 2, 2), i))
    return ops


@pytest.mark.parametrize("name", list(_circuits()))
def test_qbit_circuits_match_oracle(name):
    from syngular_amd.quantum import Circuit, Qbit

    case(f"qbit circuit {name}")
    size, structure = _circuits()[name]
    want = QO.run_circuit(size, structure)
    q = Qbit(size)
    for op in structure:
        q = q @ op
    assert rel(q.to_tensor(), want.to_tensor()) < TOL
    circ = Circuit(size, structure=list(structure))
    circ.run()
    assert rel(circ.current_state.to_tensor(), want.to_tensor()) < TOL
    assert len(circ.states) == len(structure) + 1


def test_qbit_binary_and_wide_register():
    from syngular_amd.quantum import Circuit, Qbit, gate

    case("qbit binary strings and a 30-qubit circuit")
    Qbit.LSB = False
    try:
        assert Qbit.from_binary("110").to_binary() == "110"
        assert (Qbit.from_binary("110") @ (gate.TOFFOLI, 0)).to_binary() == "111"
    finally:
        Qbit.LSB = True
    assert Qbit.from_binary("110").to_binary() == "011"
    # test/test_quantum.py:12-27 on 30 qubits: compare amplitudes and norm
    structure = [(gate.X, 0), (gate.X, 2), (gate.H, 0), (gate.H, 2)]
    circ = Circuit(size=30, structure=list(structure))
    circ.run()
    circ.add((gate.H, 1))
    circ.run()
    want = QO.run_circuit(30, [(QO.X, 0), (QO.X, 2), (QO.H, 0), (QO.H, 2), (QO.H, 1)])
    st = circ.current_state.state
    assert abs((st | st) - 1.0) < 1e-12
    for idx in [(0,) * 30, (1, 1, 1) + (0,) * 27, (0, 1, 0) + (0,) * 27, (1, 0, 1) + (0,) * 27]:
        assert abs(st[idx] - want.state[idx]) < 1e-12
