"""MPO.apply on a sub-chain (tensor/matrix_product_operator.py:582-646) on
the GPU against the oracle: the receiver's sites after the call agree with
the oracle's to a relative Frobenius error <= 1e-10 (the split is Householder
QR in the LAPACK convention, so the sites themselves match)."""
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


def mpo_sites(rng, din, dout, bonds):
    b = [1] + list(bonds) + [1]
    return [rng.standard_normal((b[k], din[k], dout[k], b[k + 1])) for k in range(len(din))]


@pytest.mark.parametrize("din,dout,bonds,odout,obonds,indices", [
    ((2, 3, 2), (3, 2, 2), (3, 4), (4,), (), [1]),
    ((2, 2, 2, 2), (2, 2, 2, 2), (3, 5, 3), (2, 2), (2,), [1, 2]),
    ((2, 2, 2, 2), (2, 2, 2, 2), (4, 16, 4), (2, 2), (3,), [1, 2]),
    ((2, 2, 2, 2, 2), (2, 2, 2, 2, 2), (4, 8, 8, 4), (2, 2, 2), (3, 3), [1, 2, 3]),
    ((4, 4, 4), (4, 4, 4), (16, 16), (4, 4), (5,), [0, 1]),
])
def test_mpo_apply_matches_oracle(din, dout, bonds, odout, obonds, indices):
    from syngular_amd.tensor import MatrixProductOperator as MPO

    case(f"mpo apply din={din} bonds={bonds} op_out={odout} indices={indices}")
    rng = np.random.default_rng(sum(din) * 10 + sum(bonds) + len(indices))
    sites = mpo_sites(rng, din, dout, bonds)
    osites = mpo_sites(rng, tuple(dout[j] for j in indices), odout, obonds)
    want = R.MPO.from_sites(list(sites))
    assert want.apply(R.MPO.from_sites(list(osites)), indices) is None
    got = MPO.from_sites(list(sites))
    assert got.apply(MPO.from_sites(list(osites)), indices) is None
    for g, w in zip(got.sites, want.sites):
        assert tuple(g.shape) == w.shape
        assert rel(g.cpu().numpy(), w) < TOL
    assert got.shape == want.shape                     # shapes are not updated


def test_mpo_apply_errors_on_device():
    from syngular_amd.tensor import MatrixProductOperator as MPO

    case("mpo apply errors")
    rng = np.random.default_rng(9)
    W = MPO.from_sites(mpo_sites(rng, (2, 2, 2), (2, 2, 2), (3, 3)))
    with pytest.raises(ValueError):
        W.apply(MPO.from_sites(mpo_sites(rng, (2, 2), (2, 2), (2,))), [0, 2])
    with pytest.raises(ValueError):
        W.apply(MPO.from_sites([rng.standard_normal((2, 2, 2, 1))]), [1])
    assert W.apply(np.eye(2), [0]) is None
