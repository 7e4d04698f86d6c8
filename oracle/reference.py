"""NumPy float64 restatement of the reference's MPS / MPO semantics.

Test infrastructure only (see oracle/__init__.py).  Each method cites the
reference lines it restates.  Behaviour that matters for parity is kept,
including the quirks the reference has:

* ``>>`` is QR truncation (projection onto the span of the leading ``dim``
  columns of each left unfolding), not an SVD
  (tensor/matrix_product_state.py:432-468).
* ``MPS >>`` returns a copy whose ``bond_shape`` is stale (computed before the
  sweep), tensor/matrix_product_state.py:433 -> :191-214.
* ``MPO >>`` first overwrites the receiver's ``bond_shape`` and then rewrites
  the receiver's site list through the aliasing copy,
  tensor/matrix_product_operator.py:477, :360, :545.
* ``@`` and ``MPO +`` finish with ``>> min_bond`` where ``min_bond`` is taken
  over the (possibly stale) bond shapes of both operands
  (tensor/matrix_product_operator.py:80,110,176,192,289).
* ``|`` is bilinear without conjugation (tensor/matrix_product_state.py:116-129);
  it is evaluated here as a streamed transfer-matrix contraction, which is the
  same sum as the reference's single opt_einsum call.

Left out on purpose: console prints (matrix_product_state.py:556,
matrix_product_operator.py:442) and the unfinished MATMUL_MODE branches
(matrix_product_operator.py:193-276).
"""
from __future__ import annotations

import itertools

import numpy as np


def _interleave_axes(n):
    # (in_0..in_{n-1}, out_0..out_{n-1}) -> (in_0, out_0, in_1, out_1, ...)
    # tensor/matrix_product_operator.py:60
    axes = []
    for i in range(n):
        axes += [i, n + i]
    return axes


def _site_shapes_mps(input_shape, bond_shape):
    # tensor/matrix_product_state.py:53-58 and :225-230
    n = len(input_shape)
    if tuple(bond_shape) == ():
        return [(1, input_shape[0], 1)]
    shapes = [(1, input_shape[0], bond_shape[0])]
    shapes += [(bond_shape[i - 1], input_shape[i], bond_shape[i]) for i in range(1, n - 1)]
    shapes += [(bond_shape[n - 2], input_shape[n - 1], 1)]
    return shapes


def _site_shapes_mpo(input_shape, output_shape, bond_shape):
    # tensor/matrix_product_operator.py:49-54 and :390-395
    n = len(input_shape)
    if tuple(bond_shape) == ():
        return [(1, input_shape[0], output_shape[0], 1)]
    shapes = [(1, input_shape[0], output_shape[0], bond_shape[0])]
    shapes += [(bond_shape[i - 1], input_shape[i], output_shape[i], bond_shape[i]) for i in range(1, n - 1)]
    shapes += [(bond_shape[n - 2], input_shape[n - 1], output_shape[n - 1], 1)]
    return shapes


class MPS:
    """Oracle MatrixProductState (tensor/matrix_product_state.py:14)."""

    def __init__(self, tensor=None, bond_shape=()):
        # tensor/matrix_product_state.py:30-61
        self.parameters_number = 0
        self.real_parameters_number = 0
        self.orthonormalized = None
        if tensor is not None:
            tensor = np.asarray(tensor, dtype=np.float64)
            self.tensor = tensor
            self.tensor_shape = tensor.shape
            self.bond_shape = bond_shape
            self.order = tensor.ndim
            self.real_parameters_number = int(np.prod(tensor.shape))
            self.input_shape = tensor.shape
            self.sites_number = len(bond_shape) + 1
            self.sites = [None] * self.sites_number
            if self.order != self.sites_number:
                raise Exception("dimensions of bond indices do not match order - 1")
            self.shape = _site_shapes_mps(self.tensor_shape, bond_shape)
        self.decomposed = False

    # ---- constructors -------------------------------------------------
    @staticmethod
    def random(input_shape, bond_shape):
        # tensor/matrix_product_state.py:175-178 (global legacy RNG)
        return MPS(np.random.normal(size=input_shape), bond_shape=bond_shape).decompose()

    @staticmethod
    def from_sites(sites, orthogonality=None, real_parameters_number=None):
        # tensor/matrix_product_state.py:191-214 (shallow list copy)
        mp = MPS()
        mp.sites = list(sites)
        mp.sites_number = len(sites)
        mp.decomposed = True
        mp.orthonormalized = orthogonality
        mp.parameters_number = int(sum(int(np.prod(s.shape)) for s in sites))
        mp.real_parameters_number = real_parameters_number
        mp.input_shape = tuple(s.shape[1] for s in sites)
        mp.bond_shape = tuple(s.shape[2] for s in sites[:-1])
        mp.shape = [tuple(s.shape) for s in sites]
        return mp

    @staticmethod
    def zeros(input_shape, bond_shape):
        # tensor/matrix_product_state.py:221-235
        return MPS.from_sites([np.zeros(s) for s in _site_shapes_mps(input_shape, bond_shape)])

    def copy(self):
        # tensor/matrix_product_state.py:282-283
        return MPS.from_sites(self.sites)

    # ---- unfoldings (tensor/matrix_product_state.py:583-622) ----------
    def _left(self, m, k):
        return np.reshape(m, (self.shape[k][1] * self.shape[k][0], -1))

    def _right(self, m, k):
        return np.reshape(m, (-1, self.shape[k][1] * self.shape[k][2]))

    def _tens(self, m, k):
        return np.reshape(m, self.shape[k])

    # ---- decomposition (tensor/matrix_product_state.py:296-350) -------
    def decompose(self, mode="left"):
        if self.decomposed:
            return self
        n = self.sites_number - 1
        T = self.tensor
        if mode == "left":
            for k in range(n):
                L = self._left(T, k)
                Q, R = np.linalg.qr(L, mode="complete")
                rank = self.shape[k][2]
                self.sites[k] = self._tens(Q[:, :rank], k)
                T = R[:rank, :]
                self.parameters_number += int(np.prod(self.sites[k].shape))
            self.sites[-1] = self._tens(T, n)
            self.parameters_number += int(np.prod(self.sites[-1].shape))
        elif mode == "right":
            for k in range(n, 0, -1):
                L = self._right(T, k).T
                Q, R = np.linalg.qr(L, mode="complete")
                rank = self.shape[k][0]
                self.sites[k] = self._tens(Q[:, :rank].T, k)
                T = R[:rank, :].T
                self.parameters_number += int(np.prod(self.sites[k].shape))
            self.sites[0] = self._tens(T, 0)
            self.parameters_number += int(np.prod(self.sites[0].shape))
        del self.tensor
        self.decomposed = True
        return self

    # ---- algebra ------------------------------------------------------
    def __add__(self, other):
        # tensor/matrix_product_state.py:75-102 (no truncation)
        if not (self.decomposed and other.decomposed):
            raise Exception("Both Matrix Product Operator must be in canonical form (use .decompose()")
        n = self.sites_number
        sites = []
        for k in range(n):
            A, B = self.sites[k], other.sites[k]
            first, last = k == 0, k == n - 1
            l = A.shape[0] if first else A.shape[0] + B.shape[0]
            r = A.shape[2] if last else A.shape[2] + B.shape[2]
            site = np.zeros((l, A.shape[1], r))
            if first:
                site[:, :, : A.shape[2]] = A
                site[:, :, A.shape[2]:] = B
            elif last:
                site[: A.shape[0]] = A
                site[A.shape[0]:] = B
            else:
                site[: A.shape[0], :, : A.shape[2]] = A
                site[A.shape[0]:, :, A.shape[2]:] = B
            sites.append(site)
        return MPS.from_sites(sites)

    def __or__(self, other):
        # tensor/matrix_product_state.py:116-129: sum over all bonds and
        # shared physical legs of prod_k X_k[a_k,s_k,a_k+1] Y_k[b_k,s_k,b_k+1]
        if not isinstance(other, MPS):
            raise Exception("right-hand site must be a MatrixProductState")
        return overlap(self.sites, other.sites)

    def dot(self):
        # tensor/matrix_product_state.py:243-244
        return np.sqrt(self | self)

    def normalize(self):
        # tensor/matrix_product_state.py:252-256
        last = self.sites[-1]
        self.sites[-1] = last / np.linalg.norm(last.reshape(last.shape[0], -1))
        return self

    def __rshift__(self, dim):
        # tensor/matrix_product_state.py:139-143
        if isinstance(dim, int):
            return self.compress(dim, strict=True)
        raise Exception("dimension should be an integer")

    def compress(self, dim, mode="left", strict=False):
        # tensor/matrix_product_state.py:362-468
        n = self.sites_number - 1
        if dim >= min(self.bond_shape):
            return self
        if strict:
            that = self.copy()
            if mode == "left":
                that._qr_truncate_sweep(dim)
            return that
        if mode == "left":
            if self.orthonormalized != "left":
                self.left_orthonormalization()
            self._qr_truncate_sweep(dim)
        elif mode == "right":
            if self.orthonormalized != "right":
                self.right_orthonormalization()
            params = 0
            for k in range(n, 0, -1):
                Rm = self._right(self.sites[k], k)
                Lm = self._left(self.sites[k - 1], k - 1)
                Q, S = np.linalg.qr(Rm.T)
                Q, S = Q[:, :dim], S[:dim, :]
                W = Lm @ S.T
                self.shape[k] = (dim,) + tuple(self.shape[k][1:])
                self.shape[k - 1] = tuple(self.shape[k - 1][:2]) + (dim,)
                self.sites[k] = self._tens(Q.T, k)
                self.sites[k - 1] = self._tens(W, k - 1)
                params += int(np.prod(self.sites[k].shape))
            self.parameters_number = params + int(np.prod(self.sites[0].shape))
        return self

    def _qr_truncate_sweep(self, dim):
        # tensor/matrix_product_state.py:439-466 (also :374-401)
        params = 0
        for k in range(self.sites_number - 1):
            L = self._left(self.sites[k], k)
            Rm = self._right(self.sites[k + 1], k + 1)
            Q, S = np.linalg.qr(L, mode="complete")
            Q, S = Q[:, :dim], S[:dim, :]
            W = S @ Rm
            self.shape[k] = tuple(self.shape[k][:2]) + (dim,)
            self.shape[k + 1] = (dim,) + tuple(self.shape[k + 1][1:])
            self.sites[k] = self._tens(Q, k)
            self.sites[k + 1] = self._tens(W, k + 1)
            params += int(np.prod(self.sites[k].shape))
        self.parameters_number = params + int(np.prod(self.sites[-1].shape))

    def left_orthonormalization(self):
        # tensor/matrix_product_state.py:554-566 (reduced QR, shapes kept)
        for k in range(self.sites_number - 1):
            L = self._left(self.sites[k], k)
            Rm = self._right(self.sites[k + 1], k + 1)
            U, B = np.linalg.qr(L)
            self.sites[k] = self._tens(U, k)
            self.sites[k + 1] = self._tens(B @ Rm, k + 1)
        self.orthonormalized = "left"

    def right_orthonormalization(self):
        # tensor/matrix_product_state.py:568-580
        for k in range(self.sites_number - 1, 0, -1):
            Rm = self._right(self.sites[k], k)
            Lm = self._left(self.sites[k - 1], k - 1)
            V, U = np.linalg.qr(Rm.T)
            self.sites[k] = self._tens(V.T, k)
            self.sites[k - 1] = self._tens(Lm @ U.T, k - 1)
        self.orthonormalized = "right"

    def left_orthogonality(self, index):
        # tensor/matrix_product_state.py:624-626
        L = self._left(self.sites[index], index)
        return L.T @ L

    def right_orthogonality(self, index):
        # tensor/matrix_product_state.py:628-630
        R = self._right(self.sites[index], index)
        return R @ R.T

    def retrieve(self, indices):
        # tensor/matrix_product_state.py:546-551
        v = np.ones((1, 1))
        for k, i in enumerate(indices):
            v = v @ self.sites[k][:, i, :]
        return v.reshape(-1)[0] if v.size == 1 else v

    def __getitem__(self, key):
        # tensor/matrix_product_state.py:157-161
        if len(key) != self.sites_number:
            raise Exception("input indices do not match the number of sites")
        return self.retrieve(key)

    def to_tensor(self):
        # tensor/matrix_product_state.py:265-273 (element-wise retrieve);
        # evaluated here by sequential contraction, same values.
        return dense_mps(self.sites)

    def apply(self, operator, index, strict=True, mode="compress"):
        """Dense m-site gate on sites index..index+m-1
        (tensor/matrix_product_state.py:487-534).

        operator has 2m legs, the first m the outputs and the last m the
        inputs contracted with the sites' physical legs (:497-501), so a
        one-site gate acts as out = G @ in.  The contracted block
        T (l, out_0..out_{m-1}, r) is split back left to right with a
        complete QR of its left unfolding, keeping the first rank columns
        of Q, rank = the site's existing right bond (:508-524): the new
        block is the projection of the old onto the span of the first rank
        columns of each unfolding.  The receiver is not modified (copy,
        :488); strict=False fails as the reference does (``that`` unbound).
        ``mode`` is unused there and here."""
        if strict:
            that = self.copy()
        m = len(operator.shape) // 2
        op = np.asarray(operator)
        if op.ndim != 2 * m:
            # opt_einsum rejects an operand with more legs than labels
            raise ValueError("operator must have as many output as input legs")
        T = that.sites[index]
        for idx in range(index + 1, index + m):
            T = np.tensordot(T, that.sites[idx], axes=([T.ndim - 1], [0]))
        # T (l, in_0..in_{m-1}, r); contract the gate's input legs
        T = np.tensordot(op, T, axes=(list(range(m, 2 * m)), list(range(1, m + 1))))
        T = np.moveaxis(T, m, 0)           # (l, out_0..out_{m-1}, r)
        for k in range(index, m + index - 1):
            lrank, input_shape = T.shape[0], T.shape[1]
            L = T.reshape(input_shape * lrank, -1)
            Q, R = np.linalg.qr(L, mode="complete")
            rank = that.shape[k][2]
            that.sites[k] = that._tens(Q[:, :rank], k)
            T = R[:rank, :].reshape((rank, -1, that.shape[k + 1][2]))
        that.sites[m + index - 1] = that._tens(T, m + index - 1)
        return that


class MPO:
    """Oracle MatrixProductOperator (tensor/matrix_product_operator.py:18)."""

    COMPRESS = True

    def __init__(self, tensor=None, bond_shape=()):
        # tensor/matrix_product_operator.py:25-64
        self.parameters_number = 0
        self.real_parameters_number = 0
        if tensor is not None:
            tensor = np.asarray(tensor, dtype=np.float64)
            self.tensor_shape = tensor.shape
            self.bond_shape = tuple(bond_shape)
            self.real_parameters_number = int(np.prod(tensor.shape))
            self.sites_number = len(self.bond_shape) + 1
            self.sites = [None] * self.sites_number
            n = self.sites_number
            self.input_shape = tuple(tensor.shape[:n])
            self.output_shape = tuple(tensor.shape[n:])
            if len(self.input_shape) != len(self.output_shape):
                raise Exception("input_shape and output_shape of the tensor must have the same length")
            if len(self.input_shape) != len(self.bond_shape) + 1:
                raise Exception("dimensions of bond indices do not match input dimension - 1")
            if self.bond_shape != ():
                self.shape = _site_shapes_mpo(self.input_shape, self.output_shape, self.bond_shape)
            else:
                self.shape = [(1, tensor.shape[0], tensor.shape[1], 1)]
            self.tensor = np.transpose(tensor, _interleave_axes(n))
        self.decomposed = False
        self.orthonormalized = None

    @staticmethod
    def random(input_shape, output_shape, bond_shape):
        # tensor/matrix_product_operator.py:352-355
        t = np.random.normal(size=(*input_shape, *output_shape))
        return MPO(t, bond_shape=bond_shape).decompose()

    @staticmethod
    def from_sites(sites, orthogonality=None, real_parameters_number=None):
        # tensor/matrix_product_operator.py:357-380 (aliases the list)
        mp = MPO()
        mp.sites = sites
        mp.sites_number = len(sites)
        mp.decomposed = True
        mp.orthonormalized = orthogonality
        mp.parameters_number = int(sum(int(np.prod(s.shape)) for s in sites))
        mp.real_parameters_number = real_parameters_number
        mp.input_shape = tuple(s.shape[1] for s in sites)
        mp.output_shape = tuple(s.shape[2] for s in sites)
        mp.bond_shape = tuple(s.shape[3] for s in sites[:-1])
        mp.shape = [tuple(s.shape) for s in sites]
        return mp

    @staticmethod
    def zeros(input_shape, output_shape, bond_shape):
        # tensor/matrix_product_operator.py:386-400
        return MPO.from_sites([np.zeros(s) for s in _site_shapes_mpo(input_shape, output_shape, bond_shape)])

    def copy(self):
        # tensor/matrix_product_operator.py:402-403
        return MPO.from_sites(self.sites)

    def _left(self, m, k):
        s = self.shape[k]
        return np.reshape(m, (s[1] * s[2] * s[0], -1))

    def _right(self, m, k):
        s = self.shape[k]
        return np.reshape(m, (-1, s[1] * s[2] * s[3]))

    def _tens(self, m, k):
        return np.reshape(m, self.shape[k])

    def decompose(self, mode="left"):
        # tensor/matrix_product_operator.py:418-467
        if self.bond_shape == ():
            self.sites = [self.tensor.reshape(1, self.tensor_shape[0], self.tensor_shape[1], 1)]
            self.decomposed = True
            del self.tensor
        if self.decomposed:
            return self
        if mode == "left":
            n = self.sites_number - 1
            T = self.tensor
            for k in range(n):
                Q, R = np.linalg.qr(self._left(T, k), mode="complete")
                rank = self.shape[k][3]
                self.sites[k] = self._tens(Q[:, :rank], k)
                T = R[:rank, :]
                self.parameters_number += int(np.prod(self.sites[k].shape))
            self.sites[-1] = self._tens(T, n)
            self.parameters_number += int(np.prod(self.sites[-1].shape))
            del self.tensor
        elif mode == "right":
            # :455-463 -- the reference decomposes left and reverses the list
            self.decompose(mode="left")
            self.sites = self.sites[::-1]
        self.decomposed = True
        return self

    def __add__(self, other):
        # tensor/matrix_product_operator.py:78-112
        min_bond = min(min(self.bond_shape), min(other.bond_shape))
        if not (self.decomposed and other.decomposed):
            raise Exception("Both Matrix Product Operator must be in canonical form (use .decompose()")
        n = self.sites_number
        sites = []
        for k in range(n):
            A, B = self.sites[k], other.sites[k]
            first, last = k == 0, k == n - 1
            l = A.shape[0] if first else A.shape[0] + B.shape[0]
            r = A.shape[3] if last else A.shape[3] + B.shape[3]
            site = np.zeros((l, A.shape[1], A.shape[2], r))
            if first:
                site[..., : A.shape[3]] = A
                site[..., A.shape[3]:] = B
            elif last:
                site[: A.shape[0]] = A
                site[A.shape[0]:] = B
            else:
                site[: A.shape[0], :, :, : A.shape[3]] = A
                site[A.shape[0]:, :, :, A.shape[3]:] = B
            sites.append(site)
        out = MPO.from_sites(sites)
        return out >> min_bond if MPO.COMPRESS else out

    def __mul__(self, mp):
        # tensor/matrix_product_operator.py:126-161: per (in, out) pair the
        # bond matrices are Kronecker-multiplied, then >> min_bond
        min_bond = min(min(self.bond_shape), min(mp.bond_shape))
        sites = []
        for k in range(self.sites_number):
            A, B = self.sites[k], mp.sites[k]
            site = np.zeros((A.shape[0] * B.shape[0], A.shape[1], A.shape[2], A.shape[3] * B.shape[3]))
            for i in range(A.shape[1]):
                for o in range(A.shape[2]):
                    site[:, i, o, :] = np.kron(A[:, i, o, :], B[:, i, o, :])
            sites.append(site)
        out = MPO.from_sites(sites)
        return out >> min_bond if MPO.COMPRESS else out

    def __matmul__(self, mp):
        # tensor/matrix_product_operator.py:175-192 (MPS) and :278-289 (MPO)
        min_bond = min(min(self.bond_shape), min(mp.bond_shape))
        sites = []
        if isinstance(mp, MPS):
            for k in range(self.sites_number):
                sites.append(apply_site(self.sites[k], mp.sites[k]))
            out = MPS.from_sites(sites)
        elif isinstance(mp, MPO):
            for k in range(self.sites_number):
                A, B = self.sites[k], mp.sites[k]
                P = np.einsum("aijb,cjpe->acipbe", A, B)
                s = self.shape[k]
                sites.append(P.reshape(s[0] * mp.shape[k][0], s[1], s[2], s[3] * mp.shape[k][3]))
            out = MPO.from_sites(sites)
        else:
            raise TypeError("unsupported operand")
        return out >> min_bond if MPO.COMPRESS else out

    def __rshift__(self, dim):
        # tensor/matrix_product_operator.py:327-331
        if isinstance(dim, int):
            return self.compress(dim, strict=True)
        raise Exception("dimension should be an integer")

    def compress(self, dim, mode="left", strict=False):
        # tensor/matrix_product_operator.py:469-580 (strict and non-strict left)
        if dim >= min(self.bond_shape):
            return self
        self.bond_shape = (dim,) * (self.sites_number - 1)
        target = self.copy() if strict else self
        if not strict and self.orthonormalized != "left":
            self.left_orthonormalization()
        params = 0
        for k in range(target.sites_number - 1):
            L = target._left(target.sites[k], k)
            Rm = target._right(target.sites[k + 1], k + 1)
            Q, S = np.linalg.qr(L, mode="complete")
            Q, S = Q[:, :dim], S[:dim, :]
            W = S @ Rm
            target.shape[k] = tuple(target.shape[k][:3]) + (dim,)
            target.shape[k + 1] = (dim,) + tuple(target.shape[k + 1][1:])
            target.sites[k] = target._tens(Q, k)
            target.sites[k + 1] = target._tens(W, k + 1)
            params += int(np.prod(target.sites[k].shape))
        target.parameters_number = params + int(np.prod(target.sites[-1].shape))
        return target

    def left_orthonormalization(self):
        # tensor/matrix_product_operator.py:673-694
        if not self.decomposed:
            raise Exception("Cannot orthonormalize an undecomposed MatrixProductOperator")
        for k in range(self.sites_number - 1):
            L = self._left(self.sites[k], k)
            Rm = self._right(self.sites[k + 1], k + 1)
            U, B = np.linalg.qr(L)
            self.sites[k] = self._tens(U, k)
            self.sites[k + 1] = self._tens(B @ Rm, k + 1)
        self.orthonormalized = "left"
        return self

    def right_orthonormalization(self):
        # tensor/matrix_product_operator.py:696-719
        if not self.decomposed:
            raise Exception("Cannot orthonormalize an undecomposed MatrixProductOperator")
        for k in range(self.sites_number - 1, 0, -1):
            Rm = self._right(self.sites[k], k)
            Lm = self._left(self.sites[k - 1], k - 1)
            V, U = np.linalg.qr(Rm.T)
            self.sites[k] = self._tens(V.T, k)
            self.sites[k - 1] = self._tens(Lm @ U.T, k - 1)
        self.orthonormalized = "right"
        return self

    def left_orthogonality(self, index):
        L = self._left(self.sites[index], index)
        return L.T @ L

    def right_orthogonality(self, index):
        R = self._right(self.sites[index], index)
        return R @ R.T

    def apply(self, operator, indices):
        """Sub-chain operator product, in place (tensor/matrix_product_operator.py:582-646).

        For an MPO operator of m = len(indices) sites on consecutive sites of
        the receiver: each receiver site's output leg is contracted with the
        operator site's input leg (:591-592), the chain is contracted into
        T (l, in_0..in_{m-1}, out_0..out_{m-1}, r) after squeezing the
        operator's outer bonds (:594-603, which must be 1), and T is split
        back left to right: reshape to (l * d_in * d_out, -1) in T's own
        memory order (so for m > 1 the rows mix in_0 with in_1, as there),
        complete QR, the first r_k columns of Q become site k with shape
        (l, d_in, d_out, r_k) and T = R[:r_k] (:605-627).  Sites are
        replaced in the receiver's list; shapes are not updated; returns
        None.  Non-consecutive indices leave labels unmatched and raise
        ValueError, as opt_einsum does.  A non-MPO operator is contracted and
        the result discarded there (:628-646); here nothing is computed."""
        if not self.decomposed:
            raise Exception("MatrixProductState not decomposed")
        if not isinstance(operator, MPO):
            return None
        m = len(indices)
        if any(indices[k + 1] != indices[k] + 1 for k in range(m - 1)):
            raise ValueError("indices must be consecutive")
        letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
        b = [next(letters) for _ in range(m + 1)]     # receiver bonds
        c = [next(letters) for _ in range(m + 1)]     # operator bonds
        i = [next(letters) for _ in range(m)]         # receiver inputs
        j = [next(letters) for _ in range(m)]         # receiver outputs = operator inputs
        p = [next(letters) for _ in range(m)]         # operator outputs
        terms, ops_ = [], []
        for k in range(m):
            terms += [b[k] + i[k] + j[k] + b[k + 1], c[k] + j[k] + p[k] + c[k + 1]]
            ops_ += [self.sites[indices[k]], operator.sites[k]]
        out = b[0] + c[0] + "".join(i) + "".join(p) + b[m] + c[m]
        T = np.einsum(",".join(terms) + "->" + out, *ops_, optimize=True)
        T = np.squeeze(T, axis=(1, T.ndim - 1))       # (l, in..., out..., r)
        for idx in range(m - 1):
            j = indices[idx]
            l, din = self.sites[j].shape[0], self.sites[j].shape[1]
            dout = operator.sites[idx].shape[2]
            L = np.reshape(T, (l * din * dout, -1))
            Q, Rr = np.linalg.qr(L, mode="complete")
            rr = self.sites[j].shape[3]
            self.sites[j] = np.reshape(Q[:, :rr], (l, din, dout, rr))
            T = Rr[:rr, :]
        j = indices[-1]
        l, din, rr = self.sites[j].shape[0], self.sites[j].shape[1], self.sites[j].shape[3]
        self.sites[j] = np.reshape(T, (l, din, operator.sites[m - 1].shape[2], rr))
        return None

    def retrieve(self, input_indices, output_indices):
        # tensor/matrix_product_operator.py:663-670
        v = np.ones((1, 1))
        for k, (i, o) in enumerate(zip(input_indices, output_indices)):
            v = v @ self.sites[k][:, i, o, :]
        return v.reshape(-1)[0]

    def __getitem__(self, key):
        # tensor/matrix_product_operator.py:333-342
        key_inp, key_out = key
        if len(key_inp) != self.sites_number:
            raise Exception("input indices do not match the number of sites")
        if len(key_out) != self.sites_number:
            raise Exception("output indices do not match the number of sites")
        return self.retrieve(key_inp, key_out)

    def to_tensor(self):
        # tensor/matrix_product_operator.py:405-416, layout (*in, *out)
        return dense_mpo(self.sites)


def mul(op1, op2):
    """syngular.mul, standard mode (tensor/utils.py:8-73)."""
    if op1.sites_number != op2.sites_number:
        raise Exception("both operator do not have the same number of sites")
    if isinstance(op1, MPS) and isinstance(op2, MPO):
        return op2 @ op1
    if isinstance(op1, MPO) and isinstance(op2, MPS):
        return op1 @ op2
    if isinstance(op1, MPS) and isinstance(op2, MPS):
        return op1 | op2
    if isinstance(op1, MPO) and isinstance(op2, MPO):
        return op2 @ op1
    raise Exception("`syn.mul` should be provided MatrixProductState or MatrixProductOperator objects only")


# ---------------------------------------------------------------------------
# site-level helpers shared with rounding.py
# ---------------------------------------------------------------------------
def apply_site(O, X):
    """P[a*Dl+c, o, b*Dr+e] = sum_s X[a,s,b] O[c,s,o,e].

    tensor/matrix_product_operator.py:184-190: contract(X,[1,2,3], O,[4,2,6,5],
    [1,4,6,3,5]) then reshape; the MPS bond index is the major one.
    """
    P = np.einsum("asb,csoe->acobe", X, O)
    return P.reshape(X.shape[0] * O.shape[0], O.shape[2], X.shape[2] * O.shape[3])


def overlap(xs, ys):
    """Bilinear <X|Y> by a left-to-right transfer sweep (no conjugation)."""
    if xs[0].shape[0] != 1 or ys[0].shape[0] != 1 or xs[-1].shape[2] != 1 or ys[-1].shape[2] != 1:
        # the reference reshapes the 4 boundary legs to a single scalar
        raise ValueError("cannot reshape array: chain boundary bonds must be 1")
    E = np.ones((1, 1))
    for X, Y in zip(xs, ys):
        E = np.einsum("ab,asc,bsd->cd", E, X, Y, optimize=True)
    return float(E[0, 0])


def dense_mps(sites):
    T = sites[0]
    for s in sites[1:]:
        T = np.tensordot(T, s, axes=([T.ndim - 1], [0]))
    return T.reshape(T.shape[1:-1])


def dense_mpo(sites):
    n = len(sites)
    T = sites[0]
    for s in sites[1:]:
        T = np.tensordot(T, s, axes=([T.ndim - 1], [0]))
    T = T.reshape(T.shape[1:-1])  # (i0,o0,i1,o1,...)
    axes = [2 * k for k in range(n)] + [2 * k + 1 for k in range(n)]
    return np.transpose(T, axes)


def retrieve_all(mp):
    """Element-by-element reconstruction, as the reference's to_tensor loops."""
    if isinstance(mp, MPS):
        out = np.zeros(mp.input_shape)
        for idx in itertools.product(*[range(d) for d in mp.input_shape]):
            out[idx] = mp.retrieve(idx)
        return out
    out = np.zeros((*mp.input_shape, *mp.output_shape))
    for i in itertools.product(*[range(d) for d in mp.input_shape]):
        for o in itertools.product(*[range(d) for d in mp.output_shape]):
            out[(*i, *o)] = mp.retrieve(i, o)
    return out
