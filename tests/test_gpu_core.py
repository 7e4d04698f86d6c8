"""Tensor-train layers on the GPU against the oracle (core/model.py:162-218):
Linear's forward is weight @ inputs with the reference's >> min_bond QR
truncation; the batched call equals the one-by-one results.  Dense outputs
agree to a relative Frobenius error <= 1e-10."""
import numpy as np
import pytest

from oracle import reference as R

pytestmark = pytest.mark.gpu

TOL = 1e-10


def case(name):
    print(f"[gpu case] {name}", flush=True)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def sites_of(mp):
    return [s.cpu().numpy() for s in mp.sites]


@pytest.mark.parametrize("units,core,bond,xbond", [(16, 2, 3, (2,)), (64, 3, 4, (3, 5)), (256, 4, 6, (4, 8, 4)),
                                                   (4096, 6, 8, (4, 8, 16, 8, 4))])
def test_linear_forward_matches_oracle(units, core, bond, xbond):
    from syngular_amd.core import Linear, Model, Output
    from syngular_amd.tensor import MatrixProductState

    case(f"linear units={units} core={core} bond={bond} x={xbond}")
    np.random.seed(units + core)
    layer = Linear(units, units, core=core, bond=bond)
    model = Model([layer, Output((units,))])
    model.build()
    W = layer.trainable_tensor_weights[0]["weight"]
    rng = np.random.default_rng(units)
    d = round(units ** (1 / core))
    b = [1] + list(xbond) + [1]
    xs = [[rng.standard_normal((b[k], d, b[k + 1])) for k in range(core)] for _ in range(5)]
    Wo = R.MPO.from_sites(sites_of(W))
    wants = [Wo @ R.MPS.from_sites(x) for x in xs]
    X = [MatrixProductState.from_sites(x) for x in xs]
    one = layer(X[0])
    assert one.shape == wants[0].shape and one.bond_shape == wants[0].bond_shape
    assert rel(one.to_tensor(), wants[0].to_tensor()) < TOL
    batch = layer(X)
    for got, want in zip(batch, wants):
        assert got.shape == want.shape and got.bond_shape == want.bond_shape
        assert got.parameters_number == want.parameters_number
        assert rel(got.to_tensor(), want.to_tensor()) < TOL
    outs = model.predict(X)
    for got, want in zip(outs, wants):
        assert rel(got, want.to_tensor().reshape(units)) < TOL


def test_model_describe_and_draw(capsys):
    from syngular_amd.core import Linear, Model, Output

    case("model describe / draw")
    np.random.seed(5)
    model = Model([Linear(64, 64, core=3, bond=4), Output((64,))])
    model.build()
    model.describe()
    out = capsys.readouterr().out
    assert "Total number of layers : 2" in out and "Layer [linear] -> linear" in out
    assert model.draw().count("O") == 3
