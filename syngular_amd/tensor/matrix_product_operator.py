"""MatrixProductOperator with device-resident cores (drop-in for
syngular.tensor.MatrixProductOperator, tensor/matrix_product_operator.py:18).

Cores are (D_l, d_in, d_out, D_r) float64 GPU tensors.  Semantics follow the
reference, quirks included: ``@`` and ``+`` finish with ``>> min_bond`` over
the (possibly stale) bond shapes of both operands; ``>>`` overwrites the
receiver's ``bond_shape`` and, because ``from_sites``/``copy`` alias the site
list, rewrites the receiver's cores; ``decompose(mode="right")`` decomposes
left and reverses the list.  ``MatrixProductOperator.COMPRESS = False``
disables the implicit ``>> min_bond`` (the reference declares the flag at
:20 but never reads it).  The unfinished MATMUL_MODE branches (:193-276)
raise NotImplementedError instead of returning None.
"""
from __future__ import annotations

from typing import List, Tuple, Union

import numpy as np
import torch

from .. import linalg, ops
from .matrix_product_state import MatrixProductState, to_device

F64 = torch.float64


def site_shapes(input_shape, output_shape, bond_shape):
    # tensor/matrix_product_operator.py:49-54, :390-395
    n = len(input_shape)
    if tuple(bond_shape) == ():
        return [(1, input_shape[0], output_shape[0], 1)]
    shape = [(1, input_shape[0], output_shape[0], bond_shape[0])]
    shape += [(bond_shape[i - 1], input_shape[i], output_shape[i], bond_shape[i]) for i in range(1, n - 1)]
    shape += [(bond_shape[n - 2], input_shape[n - 1], output_shape[n - 1], 1)]
    return shape


class MatrixProductOperator:
    COMPRESS = True
    DECOMPOSE = False
    MATMUL_MODE = "standard"

    def __init__(self, tensor=None, bond_shape: Union[Tuple, List] = (), verbose=0) -> None:
        # tensor/matrix_product_operator.py:25-64
        self.parameters_number = 0
        self.real_parameters_number = 0
        if tensor is not None:
            t = to_device(tensor)
            self.tensor_shape = tuple(t.shape)
            self.bond_shape = tuple(bond_shape)
            self.real_parameters_number = int(np.prod(self.tensor_shape))
            self.sites_number = len(self.bond_shape) + 1
            self.sites = [None] * self.sites_number
            n = self.sites_number
            self.input_shape = tuple(self.tensor_shape[:n])
            self.output_shape = tuple(self.tensor_shape[n:])
            if len(self.input_shape) != len(self.output_shape):
                raise Exception("input_shape and output_shape of the tensor must have the same length")
            if len(self.input_shape) != len(self.bond_shape) + 1:
                raise Exception("dimensions of bond indices do not match input dimension - 1")
            if self.bond_shape != ():
                self.shape = site_shapes(self.input_shape, self.output_shape, self.bond_shape)
            else:
                self.shape = [(1, self.tensor_shape[0], self.tensor_shape[1], 1)]
            self.shape_indices = [(2 * n + i, i, n + i, 2 * n + i + 1) for i in range(n)]
            # interleave (in_0..in_n-1, out_0..out_n-1) -> (in_0, out_0, in_1, out_1, ...)
            axes = []
            for i in range(n):
                axes += [i, n + i]
            self.tensor = t.permute(axes).contiguous()
        self.verbose = verbose
        self.decomposed = False
        self.orthonormalized = None

    # ------------------------------------------------------------------
    @staticmethod
    def random(input_shape, output_shape, bond_shape):
        # tensor/matrix_product_operator.py:352-355 (global legacy NumPy RNG)
        tensor = np.random.normal(size=(*input_shape, *output_shape))
        return MatrixProductOperator(tensor, bond_shape=bond_shape).decompose()

    @staticmethod
    def from_sites(sites, orthogonality=None, real_parameters_number=None):
        # tensor/matrix_product_operator.py:357-380 -- aliases the list
        for i, s in enumerate(sites):
            if not (isinstance(s, torch.Tensor) and s.is_cuda and s.dtype == F64 and s.is_contiguous()):
                sites[i] = to_device(s)
        mp = MatrixProductOperator()
        mp.sites = sites
        mp.sites_number = len(sites)
        mp.decomposed = True
        mp.orthonormalized = orthogonality
        mp.parameters_number = int(sum(s.numel() for s in sites))
        mp.real_parameters_number = real_parameters_number
        mp.input_shape = tuple(int(s.shape[1]) for s in sites)
        mp.output_shape = tuple(int(s.shape[2]) for s in sites)
        mp.bond_shape = tuple(int(s.shape[3]) for s in sites[:-1])
        mp.shape = [tuple(int(x) for x in s.shape) for s in sites]
        return mp

    @staticmethod
    def empty():
        return MatrixProductOperator()

    @staticmethod
    def zeros(input_shape, output_shape, bond_shape):
        # tensor/matrix_product_operator.py:386-400
        dev = ops.device()
        return MatrixProductOperator.from_sites(
            [torch.zeros(s, dtype=F64, device=dev) for s in site_shapes(input_shape, output_shape, bond_shape)])

    def copy(self):
        # tensor/matrix_product_operator.py:402-403
        return MatrixProductOperator.from_sites(self.sites)

    def decompose(self, mode="left"):
        # tensor/matrix_product_operator.py:418-467
        if self.bond_shape == ():
            self.sites = [linalg.reshape(self.tensor, (1, self.tensor_shape[0], self.tensor_shape[1], 1))]
            self.decomposed = True
            del self.tensor
        if not self.decomposed:
            if mode == "left":
                self.sites, params = linalg.decompose_left(self.tensor, self.shape)
                self.parameters_number += params
                del self.tensor
            elif mode == "right":
                self.decompose(mode="left")
                self.sites = self.sites[::-1]
            self.decomposed = True
        return self

    # ------------------------------------------------------------------
    def _finish(self, out, min_bond):
        return out >> min_bond if MatrixProductOperator.COMPRESS else out

    def __add__(self, mpo):
        # tensor/matrix_product_operator.py:78-112
        min_bond = min(min(self.bond_shape), min(mpo.bond_shape))
        if not (self.decomposed and mpo.decomposed):
            raise Exception("Both Matrix Product Operator must be in canonical form (use .decompose()")
        n = self.sites_number
        triples = [(self.sites[k], mpo.sites[k], 1 if k == 0 else (2 if k == n - 1 else 0)) for k in range(n)]
        return self._finish(MatrixProductOperator.from_sites(ops.direct_sum_many(triples)), min_bond)

    def __mul__(self, mp):
        # tensor/matrix_product_operator.py:126-161: core-wise Kronecker
        # product of the bond legs for each (in, out) pair (Hadamard product
        # of the two operators), then >> min_bond.
        min_bond = min(min(self.bond_shape), min(mp.bond_shape))
        if not self.decomposed and mp.decomposed:
            raise Exception("Both Matrix Product Operator must be in canonical form (use .decompose()")
        if isinstance(mp, MatrixProductOperator):
            sites = []
            for k in range(self.sites_number):
                A, B = self.sites[k], mp.sites[k]
                sites.append(_kron_bonds(A, B))
            return self._finish(MatrixProductOperator.from_sites(sites), min_bond)
        if isinstance(mp, MatrixProductState):
            return None
        raise Exception("left hand-side must be either a MatrixProductState or a MatrixProductOperator")

    def __matmul__(self, mp):
        # tensor/matrix_product_operator.py:175-289
        min_bond = min(min(self.bond_shape), min(mp.bond_shape))
        if isinstance(mp, MatrixProductState):
            if MatrixProductOperator.MATMUL_MODE != "standard":
                raise NotImplementedError(
                    f"MATMUL_MODE={MatrixProductOperator.MATMUL_MODE!r} is unfinished in the reference")
            sites = ops.apply_many(list(zip(self.sites, mp.sites)))
            out = MatrixProductState.from_sites(sites)
            return out >> min_bond if MatrixProductOperator.COMPRESS else out
        if isinstance(mp, MatrixProductOperator):
            sites = ops.mpo_mpo_many(list(zip(self.sites, mp.sites)))
            # the reference reshapes with self.shape[idx][2] as the output leg
            for k, s in enumerate(sites):
                sh = self.shape[k]
                sites[k] = linalg.reshape(s, (sh[0] * mp.shape[k][0], sh[1], sh[2], sh[3] * mp.shape[k][3]))
            return self._finish(MatrixProductOperator.from_sites(sites), min_bond)
        return None

    def __mod__(self, mode):
        # tensor/matrix_product_operator.py:291-315 (a stub in the reference)
        if mode in ("L", "left", "l", "left-canonical", "R", "right", "r", "right-canonical"):
            return None
        raise Exception("when calling mod (%) mode should be left-canonical or right-canonical alias")

    def __rshift__(self, dim: int):
        # tensor/matrix_product_operator.py:327-331
        if isinstance(dim, int):
            return self.compress(dim, strict=True)
        raise Exception("dimension should be an integer")

    def __getitem__(self, key):
        # tensor/matrix_product_operator.py:333-342
        key_inp, key_out = key[0], key[1]
        if len(key_inp) != self.sites_number:
            raise Exception("input indices do not match the number of sites")
        if len(key_out) != self.sites_number:
            raise Exception("output indices do not match the number of sites")
        return self.retrieve(key_inp, key_out)

    def __repr__(self) -> str:
        r = "<Matrix Product Operator> \n> Sites shape" + str(self.shape) + "\n"
        r += "\t" + "|   " * self.sites_number + "\n"
        r += "\t" + ("O---" * (self.sites_number - 1)) + "O" + "\n"
        r += "\t" + "|   " * self.sites_number + "\n"
        return r

    # ------------------------------------------------------------------
    def compress(self, dim, mode="left", strict=False):
        # tensor/matrix_product_operator.py:469-580
        if dim >= min(self.bond_shape):
            return self
        self.bond_shape = (dim,) * (self.sites_number - 1)
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
            params = 0
            for k in range(self.sites_number - 1, 0, -1):
                Rm = linalg.right_unfold(self.sites[k], self.shape[k])
                Lm = linalg.left_unfold(self.sites[k - 1], self.shape[k - 1])
                Qt = ops.householder_q(Rm.t(), min(min(Rm.shape), dim), q_trans=True)
                St = ops.gemm(Rm, Qt.t())
                self.shape[k] = (dim,) + tuple(self.shape[k][1:])
                self.shape[k - 1] = tuple(self.shape[k - 1][:-1]) + (dim,)
                self.sites[k] = linalg.reshape(Qt, self.shape[k])
                self.sites[k - 1] = linalg.reshape(ops.gemm(Lm, St), self.shape[k - 1])
                params += self.sites[k].numel()
            self.parameters_number = params + self.sites[0].numel()
        return self

    def left_orthonormalization(self, bond_shape=()):
        # tensor/matrix_product_operator.py:673-694
        if not self.decomposed:
            raise Exception("Cannot orthonormalize an undecomposed MatrixProductOperator")
        linalg.left_orthonormalize(self.sites, self.shape)
        self.orthonormalized = "left"
        return self

    def right_orthonormalization(self, bond_shape=()):
        # tensor/matrix_product_operator.py:696-719
        if not self.decomposed:
            raise Exception("Cannot orthonormalize an undecomposed MatrixProductOperator")
        linalg.right_orthonormalize(self.sites, self.shape)
        self.orthonormalized = "right"
        return self

    def left_orthogonality(self, index):
        L = linalg.left_unfold(self.sites[index], self.shape[index])
        return ops.gemm(L.t(), L).cpu().numpy()

    def right_orthogonality(self, index):
        R = linalg.right_unfold(self.sites[index], self.shape[index])
        return ops.gemm(R, R.t()).cpu().numpy()

    def retrieve(self, input_indices, output_indices):
        # tensor/matrix_product_operator.py:663-670
        mats = [self.sites[k][:, int(i), int(o), :] for k, (i, o) in enumerate(zip(input_indices, output_indices))]
        v = linalg.row_product(mats)
        return float(v.reshape(-1)[0].item()) if v.numel() == 1 else v.cpu().numpy()

    def to_tensor(self):
        # tensor/matrix_product_operator.py:405-416, layout (*in, *out)
        n = self.sites_number
        d = linalg.dense(self.sites)
        d = d.reshape([s for pair in zip(self.input_shape, self.output_shape) for s in pair])
        axes = [2 * k for k in range(n)] + [2 * k + 1 for k in range(n)]
        return d.permute(axes).cpu().numpy()

    def left_site_matricization(self, index):
        return linalg.left_unfold(self.sites[index], self.shape[index])

    def right_site_matricization(self, index):
        return linalg.right_unfold(self.sites[index], self.shape[index])

    def tensoricization(self, matrix, index):
        return linalg.reshape(matrix, self.shape[index])

    # ------------------------------------------------------------------
    def apply(self, operator, indices):
        """Sub-chain operator product, in place (tensor/matrix_product_operator.py:582-646).

        operator (an MPO of m = len(indices) sites) acts on the consecutive
        receiver sites `indices`: each receiver site's output leg meets the
        operator site's input leg.  On the device:

        1. P_k = receiver site x operator site (the MPO x MPO site kernel),
           bonds (receiver, operator) pairs;
        2. the chain of P_k is contracted into T (l, in_0..in_{m-1},
           out_0..out_{m-1}, r), the reference's leg order (:594-603), one
           grouped GEMM launch per site: block (l, ins) x in_k of the new T
           is  T_k[(l, ins), outs, :] @ P_k[:, in_k, :, :];  the operator's
           outer bonds must be 1 (the reference squeezes them);
        3. the split of :605-627: T reshaped to (l d_in d_out, -1) in its own
           memory order (for m > 1 the rows mix in_0 with in_1, as there),
           Householder QR, the first r_k columns of Q as site k
           (l, d_in, d_out, r_k) and T = Q^T L.

        The receiver's site list is rewritten; shapes are not updated;
        returns None.  Non-consecutive indices raise ValueError (unmatched
        labels in the reference's contraction).  A non-MPO operator is
        contracted and discarded by the reference (:628-646); here nothing is
        computed."""
        if not self.decomposed:
            raise Exception("MatrixProductState not decomposed")
        if not isinstance(operator, MatrixProductOperator):
            return None
        m = len(indices)
        if any(indices[k + 1] != indices[k] + 1 for k in range(m - 1)):
            raise ValueError("indices must be consecutive")
        P = ops.mpo_mpo_many([(self.sites[j], operator.sites[k]) for k, j in enumerate(indices)])
        squeeze = "cannot select an axis to squeeze out which has size not equal to one"
        if operator.sites[0].shape[0] != 1 or operator.sites[m - 1].shape[3] != 1:
            raise ValueError(squeeze)
        l = int(P[0].shape[0])
        T = P[0]                                           # (l, i0, p0, B1), grouped for m = 1
        n_in, n_out = int(T.shape[1]), int(T.shape[2])     # products of the ins / outs so far
        for k in range(1, m):
            Pk = P[k]
            bk, ik, pk, bn = (int(x) for x in Pk.shape)
            Tk = T.reshape(l * n_in, n_out, bk)
            Tn = torch.empty((l * n_in, ik, n_out, pk * bn), dtype=torch.float64, device=T.device)
            Pv = Pk.reshape(bk, ik, pk * bn)
            ops.gemm_many([(Tk[q], Pv[:, a, :], Tn[q, a], 1.0, 0.0) for q in range(l * n_in) for a in range(ik)])
            T = Tn
            n_in, n_out = n_in * ik, n_out * pk
        T = T.reshape(l, n_in, n_out, -1)                  # (l, ins, outs, r) with the unit operator bond
        for idx in range(m - 1):
            j = indices[idx]
            sl, din = int(self.sites[j].shape[0]), int(self.sites[j].shape[1])
            dout = int(operator.sites[idx].shape[2])
            rows = sl * din * dout
            if rows == 0 or T.numel() % rows:
                raise ValueError(f"cannot reshape array of size {T.numel()} into shape ({rows},newaxis)")
            L = T.contiguous().reshape(rows, -1)
            rr = int(self.sites[j].shape[3])
            Q, S = linalg.qr_keep(L, rr)
            self.sites[j] = linalg.reshape(Q, (sl, din, dout, rr))
            T = S
        j = indices[-1]
        sl, din, rr = (int(self.sites[j].shape[x]) for x in (0, 1, 3))
        self.sites[j] = linalg.reshape(T, (sl, din, int(operator.sites[m - 1].shape[2]), rr))
        return None


def _kron_bonds(A, B):
    """site[:, i, o, :] = kron(A[:, i, o, :], B[:, i, o, :]) for every (i, o)
    (tensor/matrix_product_operator.py:140-152), as one MPO x MPO-style
    outer product over the bond legs with the physical legs matched."""
    al, din, dout, ar = A.shape
    bl, din2, dout2, br = B.shape
    if (din, dout) != (din2, dout2):
        raise ValueError("could not broadcast: physical dimensions differ")
    return ops.kron_bonds(A, B)
