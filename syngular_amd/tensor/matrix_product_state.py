"""MatrixProductState with device-resident cores (drop-in for
syngular.tensor.MatrixProductState, tensor/matrix_product_state.py:14).

Attributes, method names, argument meaning and error behaviour follow the
reference, including its quirks (stale ``bond_shape`` on the copy that ``>>``
returns, no truncation in ``+``).  What changes:

* ``sites`` is a list of float64 torch tensors on the GPU; every arithmetic
  step runs in the gfx950 kernels of libsyngular_amd.so.
* results that the reference returns as small NumPy values stay NumPy-shaped
  on the host: ``|`` and ``dot`` return floats, ``[]``/``retrieve`` return a
  float, ``to_tensor`` and the orthogonality checks return NumPy arrays.
* console prints of the reference (matrix_product_state.py:556) are dropped.
* extension: ``round(chi_max, cutoff)`` is the SVD rounding; with
  ``MatrixProductState.ROUNDING = "svd"`` the ``>>`` operator uses it too
  (``>> chi`` = ``round(chi_max=chi, cutoff=MatrixProductState.CUTOFF)``).
"""
from __future__ import annotations

import math
from typing import List, Tuple, Union

import numpy as np
import torch

from .. import linalg, ops, rounding

F64 = torch.float64


def to_device(t):
    """Accept numpy arrays, Python sequences or tensors; return a float64
    contiguous tensor on the current GPU."""
    if isinstance(t, torch.Tensor):
        return t.to(device=ops.device(), dtype=F64).contiguous()
    return torch.as_tensor(np.asarray(t, dtype=np.float64), device=ops.device()).contiguous()


def gate_matrix(operator, m):
    """The (D_out, D_in) float64 device matrix of a gate tensor with m output
    legs followed by m input legs.  A complex gate is accepted only when its
    imaginary part is zero."""
    if len(operator.shape) != 2 * m:
        raise ValueError(f"a gate has as many output as input legs, got shape {tuple(operator.shape)}")
    if isinstance(operator, torch.Tensor):
        if operator.is_complex():
            if bool((operator.imag != 0).any()):
                raise TypeError("complex gates are not supported: the device path computes in float64")
            operator = operator.real
    else:
        operator = np.asarray(operator)
        if np.iscomplexobj(operator):
            if np.any(operator.imag != 0):
                raise TypeError("complex gates are not supported: the device path computes in float64")
            operator = operator.real
    dout = math.prod(int(s) for s in operator.shape[:m])
    din = math.prod(int(s) for s in operator.shape[m:])
    return to_device(operator).reshape(dout, din)


def site_shapes(input_shape, bond_shape):
    # tensor/matrix_product_state.py:53-58, :225-230
    n = len(input_shape)
    if tuple(bond_shape) == ():
        return [(1, input_shape[0], 1)]
    shape = [(1, input_shape[0], bond_shape[0])]
    shape += [(bond_shape[i - 1], input_shape[i], bond_shape[i]) for i in range(1, n - 1)]
    shape += [(bond_shape[n - 2], input_shape[n - 1], 1)]
    return shape


class MatrixProductState:
    ROUNDING = "qr"    # ">>" semantics: "qr" (reference) or "svd"
    CUTOFF = 0.0       # relative discarded-weight cutoff used by ">>" in "svd" mode

    def __init__(self, tensor=None, bond_shape: Union[Tuple, List] = (), verbose=0) -> None:
        # tensor/matrix_product_state.py:30-61
        self.parameters_number = 0
        self.real_parameters_number = 0
        self.orthonormalized = None
        if tensor is not None:
            self.tensor = to_device(tensor)
            self.tensor_shape = tuple(self.tensor.shape)
            self.bond_shape = bond_shape
            self.order = len(self.tensor_shape)
            self.real_parameters_number = int(np.prod(self.tensor_shape))
            self.input_shape = self.tensor_shape
            self.sites_number = len(self.bond_shape) + 1
            self.sites = [None] * self.sites_number
            self.norm = None
            if self.order != self.sites_number:
                raise Exception("dimensions of bond indices do not match order - 1")
            self.shape = site_shapes(self.tensor_shape, bond_shape)
        self.verbose = verbose
        self.decomposed = False

    # ------------------------------------------------------------------
    # construction
    # ------------------------------------------------------------------
    @staticmethod
    def random(input_shape, bond_shape):
        """tensor/matrix_product_state.py:175-178: a dense N(0,1) tensor drawn
        from NumPy's global legacy generator (same values as the reference
        for the same np.random.seed), decomposed on the device."""
        tensor = np.random.normal(size=input_shape)
        return MatrixProductState(tensor, bond_shape=bond_shape).decompose()

    @staticmethod
    def from_sites(sites, orthogonality=None, real_parameters_number=None):
        # tensor/matrix_product_state.py:191-214 (shallow list copy)
        mp = MatrixProductState()
        mp.sites = [to_device(s) for s in sites]
        mp.sites_number = len(sites)
        mp.decomposed = True
        mp.orthonormalized = orthogonality
        mp.parameters_number = int(sum(s.numel() for s in mp.sites))
        mp.real_parameters_number = real_parameters_number
        mp.input_shape = tuple(int(s.shape[1]) for s in mp.sites)
        mp.bond_shape = tuple(int(s.shape[2]) for s in mp.sites[:-1])
        mp.shape = [tuple(int(x) for x in s.shape) for s in mp.sites]
        return mp

    @staticmethod
    def empty():
        return MatrixProductState()

    @staticmethod
    def zeros(input_shape, bond_shape):
        # tensor/matrix_product_state.py:221-235
        dev = ops.device()
        return MatrixProductState.from_sites(
            [torch.zeros(s, dtype=F64, device=dev) for s in site_shapes(input_shape, bond_shape)])

    def copy(self):
        # tensor/matrix_product_state.py:282-283 (cores are shared)
        return MatrixProductState.from_sites(self.sites)

    def decompose(self, mode="left"):
        # tensor/matrix_product_state.py:296-350
        if not self.decomposed:
            if mode == "left":
                self.sites, params = linalg.decompose_left(self.tensor, self.shape)
            elif mode == "right":
                self.sites, params = linalg.decompose_right(self.tensor, self.shape)
            else:
                return self
            self.parameters_number += params
            del self.tensor
            self.decomposed = True
        return self

    # ------------------------------------------------------------------
    # operators
    # ------------------------------------------------------------------
    def __add__(self, mps):
        # tensor/matrix_product_state.py:75-102 (direct sum, no truncation)
        if not (self.decomposed and mps.decomposed):
            raise Exception("Both Matrix Product Operator must be in canonical form (use .decompose()")
        n = self.sites_number
        triples = []
        for k in range(n):
            edge = 1 if k == 0 else (2 if k == n - 1 else 0)
            if n == 1:
                raise ValueError("could not broadcast input array: a single-site direct sum has no bond leg")
            triples.append((self.sites[k], mps.sites[k], edge))
        return MatrixProductState.from_sites(ops.direct_sum_many(triples))

    def __or__(self, mp):
        # tensor/matrix_product_state.py:116-129
        if isinstance(mp, MatrixProductState):
            return float(linalg.overlap(self.sites, mp.sites).item())
        raise Exception("right-hand site must be a MatrixProductState")

    def __rshift__(self, dim: int):
        # tensor/matrix_product_state.py:139-143
        if isinstance(dim, int):
            if MatrixProductState.ROUNDING == "svd":
                return self.round(chi_max=dim, cutoff=MatrixProductState.CUTOFF)
            return self.compress(dim, strict=True)
        raise Exception("dimension should be an integer")

    def __getitem__(self, key):
        # tensor/matrix_product_state.py:157-161
        if len(key) != self.sites_number:
            raise Exception("input indices do not match the number of sites")
        return self.retrieve(key)

    def __repr__(self) -> str:
        r = "<Matrix Product State> \n> Sites shape" + str(self.shape) + "\n"
        r += "\t" + "|   " * self.sites_number + "\n"
        r += "\t" + ("O---" * (self.sites_number - 1)) + "O" + "\n"
        return r

    # ------------------------------------------------------------------
    # compression
    # ------------------------------------------------------------------
    def compress(self, dim: int, mode="left", strict=False):
        # tensor/matrix_product_state.py:362-468
        if dim >= min(self.bond_shape):
            return self
        if strict:
            that = self.copy()
            if mode == "left":
                that.parameters_number = linalg.qr_truncate_sweep(that.sites, that.shape, dim)
            return that
        if mode == "left":
            if self.orthonormalized != "left":
                self.left_orthonormalization()
            self.parameters_number = linalg.qr_truncate_sweep(self.sites, self.shape, dim)
        elif mode == "right":
            if self.orthonormalized != "right":
                self.right_orthonormalization()
            self._right_qr_truncate(dim)
        return self

    def _right_qr_truncate(self, dim):
        # tensor/matrix_product_state.py:406-430: Q, S = qr(R.T) (reduced),
        # keep Q[:, :dim], S[:dim]; site k = Q.T, site k-1 = L @ S.T
        params = 0
        for k in range(self.sites_number - 1, 0, -1):
            Rm = linalg.right_unfold(self.sites[k], self.shape[k])
            Lm = linalg.left_unfold(self.sites[k - 1], self.shape[k - 1])
            red = min(Rm.shape)
            Qt = ops.householder_q(Rm.t(), min(red, dim), q_trans=True)    # Q[:, :dim].T
            St = ops.gemm(Rm, Qt.t())                                       # S[:dim].T = R Q
            self.shape[k] = (dim,) + tuple(self.shape[k][1:])
            self.shape[k - 1] = tuple(self.shape[k - 1][:-1]) + (dim,)
            self.sites[k] = linalg.reshape(Qt, self.shape[k])
            self.sites[k - 1] = linalg.reshape(ops.gemm(Lm, St), self.shape[k - 1])
            params += self.sites[k].numel()
        self.parameters_number = params + self.sites[0].numel()

    def round(self, chi_max=None, cutoff=0.0, report=None):
        """SVD rounding (left QR sweep, right Jacobi-SVD sweep) on the device.
        Returns a new MPS; pass a rounding.RoundingReport to receive the
        per-bond singular values, kept ranks and discarded weights."""
        cores = rounding.tt_round(self.sites, chi_max, cutoff, report)
        return MatrixProductState.from_sites(cores, orthogonality="right")

    # ------------------------------------------------------------------
    # dense gates
    # ------------------------------------------------------------------
    def apply(self, operator, index, strict=True, mode="compress"):
        """Dense m-site gate on sites index..index+m-1
        (tensor/matrix_product_state.py:487-534).

        ``operator`` (NumPy or torch, real) has 2m legs: the first m are the
        outputs, the last m the inputs contracted with the sites' physical
        legs, so a one-site gate acts as out = G @ in.  On the device:

        1. the m sites are multiplied into one block T (l, D_in, r) by the
           GEMM, D_in = the product of their physical dimensions;
        2. the gate goes through the MPO x MPS apply kernel as a bond-1 MPO
           core O[0, s, o, 0] = G[o, s], giving T (l, D_out, r);
        3. the block is split back left to right as the reference does: the
           left unfolding L of T is factored by Householder QR (the LAPACK
           reflector convention, so zero columns give H = I as in numpy),
           site k takes the first ``rank`` columns of the complete Q, rank =
           its existing right bond, and T = Q_rank^T L (the leading rows of
           R), reshaped (rank, -1, right bond of site k+1) exactly as the
           reference reshapes it.

        The receiver is not modified.  strict=False fails as the reference
        does (``that`` is unbound, UnboundLocalError); ``mode`` is unused
        there and here.  Shapes that the reference cannot reshape raise
        ValueError.  Complex gates with a nonzero imaginary part raise
        TypeError: the device path computes in float64."""
        if strict:
            that = self.copy()
        m = len(operator.shape) // 2
        G = gate_matrix(operator, m)
        T = that.sites[index]
        for idx in range(index + 1, index + m):
            nxt = that.sites[idx]
            T = ops.gemm(T.reshape(-1, T.shape[-1]), nxt.reshape(nxt.shape[0], -1))
            T = T.reshape(that.sites[index].shape[0], -1, nxt.shape[-1])
        if T.shape[1] != G.shape[1]:
            raise ValueError(f"operands could not be broadcast: gate input dimension {G.shape[1]} "
                             f"vs sites {T.shape[1]}")
        O = G.t().contiguous().reshape(1, G.shape[1], G.shape[0], 1)
        T = ops.apply_many([(O, T)])[0]
        outs = tuple(int(s) for s in operator.shape[:m])
        T = T.reshape(T.shape[0], *outs, T.shape[-1])
        for k in range(index, m + index - 1):
            lrank, input_shape = T.shape[0], T.shape[1]
            L = linalg.reshape(T, (input_shape * lrank, T.numel() // (input_shape * lrank)))
            rank = that.shape[k][2]
            Q = ops.householder_q(L, rank)                  # complete Q[:, :rank]
            that.sites[k] = that.tensoricization(Q, k)
            R = ops.gemm(Q.t(), L)                          # R[:rank] = Q_rank^T L
            nr = that.shape[k + 1][2]
            if R.numel() % (rank * nr):
                raise ValueError(f"cannot reshape array of size {R.numel()} into shape ({rank},newaxis,{nr})")
            T = R.reshape(rank, -1, nr)
        that.sites[m + index - 1] = that.tensoricization(T, m + index - 1)
        return that

    # ------------------------------------------------------------------
    # orthonormalisation
    # ------------------------------------------------------------------
    def left_orthonormalization(self, bond_shape=()):
        # tensor/matrix_product_state.py:554-566
        linalg.left_orthonormalize(self.sites, self.shape)
        self.orthonormalized = "left"

    def right_orthonormalization(self, bond_shape=()):
        # tensor/matrix_product_state.py:568-580
        linalg.right_orthonormalize(self.sites, self.shape)
        self.orthonormalized = "right"

    def left_orthogonality(self, index):
        # tensor/matrix_product_state.py:624-626: L.T @ L
        L = linalg.left_unfold(self.sites[index], self.shape[index])
        return ops.gemm(L.t(), L).cpu().numpy()

    def right_orthogonality(self, index):
        # tensor/matrix_product_state.py:628-630: R @ R.T
        R = linalg.right_unfold(self.sites[index], self.shape[index])
        return ops.gemm(R, R.t()).cpu().numpy()

    # ------------------------------------------------------------------
    # scalars and dense views
    # ------------------------------------------------------------------
    def dot(self):
        # tensor/matrix_product_state.py:243-244
        return float(np.sqrt(self | self))

    def normalize(self):
        # tensor/matrix_product_state.py:252-256: last core /= its Frobenius norm
        last = self.sites[-1]
        nrm = math.sqrt(float(linalg.frob_norm2(last).item()))
        s = torch.full((1,), nrm, dtype=F64, device=last.device)
        flat = last.contiguous().reshape(1, -1)
        ops.rowscale_many([(flat, s, 1)])
        self.sites[-1] = flat.reshape(last.shape)
        return self

    def retrieve(self, indices):
        # tensor/matrix_product_state.py:546-551
        mats = [self.sites[k][:, int(i), :] for k, i in enumerate(indices)]
        v = linalg.row_product(mats)
        return float(v.reshape(-1)[0].item()) if v.numel() == 1 else v.cpu().numpy()

    def to_tensor(self):
        # tensor/matrix_product_state.py:265-273 (dense reconstruction)
        return linalg.dense(self.sites).reshape(self.input_shape).cpu().numpy()

    def grad(self, index):
        return self.sites[:index] + self.sites[index + 2:]

    # matricisation helpers with the reference names (:583-622)
    def left_site_matricization(self, index):
        return linalg.left_unfold(self.sites[index], self.shape[index])

    def right_site_matricization(self, index):
        return linalg.right_unfold(self.sites[index], self.shape[index])

    def tensoricization(self, matrix, index):
        return linalg.reshape(matrix, self.shape[index])
